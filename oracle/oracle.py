"""ctypes front end of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py use this module, and only to check results against. It wraps
``oracle/_build/libbm3d_oracle.so`` (built from ``bm3d_oracle.c`` by
``oracle/build_oracle.py``) and restates the host-side pieces of the
reference plugin that sit around the kernel:

* frame-window assembly of ``BM3DGetFrame`` (source/source.cpp:206-264):
  slot z of frame n is frame clamp(n - r + z, 0, N - 1);
* the r = 0 division ``Aggregation`` (source/source.cpp:170-188);
* ``VAggregateGetFrame`` (source/source.cpp:746-862);
* the composition of ``BM3Dv2Create`` (source/source.cpp:911-989);
* sigma scaling of ``BM3DCreate`` (source/source.cpp:520-523).

Clips are numpy arrays shaped [frames, planes, height, width].
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libbm3d_oracle.so")
_lib = None

FLT_EPSILON = np.float32(np.finfo(np.float32).eps)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(
                f"oracle library missing: {_LIB_PATH} (run oracle/build_oracle.py)")
        L = ctypes.CDLL(_LIB_PATH)
        for sfx, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            P = ctypes.POINTER(ct)
            PP = ctypes.POINTER(P)
            bm3d_args = [P, PP, PP, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                         ctypes.POINTER(ctypes.c_float), ctypes.c_int, ctypes.c_int,
                         ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                         ctypes.c_int, ctypes.c_float, ctypes.POINTER(ctypes.c_int)]
            f = getattr(L, f"oracle_bm3d_{sfx}")
            f.restype = ctypes.c_longlong
            f.argtypes = bm3d_args
            f = getattr(L, f"oracle_bm3d_ex_{sfx}")
            f.restype = ctypes.c_longlong
            f.argtypes = bm3d_args + [ctypes.c_int] * 3
            match_args = [PP, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                          ctypes.POINTER(ctypes.c_int), P]
            f = getattr(L, f"oracle_match_{sfx}")
            f.restype = ctypes.c_longlong
            f.argtypes = match_args
            f = getattr(L, f"oracle_match_ex_{sfx}")
            f.restype = ctypes.c_longlong
            f.argtypes = match_args + [ctypes.c_int]
            f = getattr(L, f"oracle_xform8_{sfx}")
            f.restype = None
            f.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"oracle_transform3d_kinds_{sfx}")
            f.restype = None
            f.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"oracle_block_cost_{sfx}")
            f.restype = ct
            f.argtypes = [P, P, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"oracle_dct8_{sfx}")
            f.restype = None
            f.argtypes = [P, ctypes.c_int, ctypes.c_int]
            f = getattr(L, f"oracle_transform3d_{sfx}")
            f.restype = None
            f.argtypes = [P, ctypes.c_int]
            f = getattr(L, f"oracle_ssd_{sfx}")
            f.restype = ct
            f.argtypes = [P, P, ctypes.c_int]
        for sfx, ct in (("f32", ctypes.c_float), ("f64", ctypes.c_double)):
            P = ctypes.POINTER(ct)
            f = getattr(L, f"oracle_collab_hard_{sfx}")
            f.restype = ct
            f.argtypes = [P, ct]
            f = getattr(L, f"oracle_collab_wiener_{sfx}")
            f.restype = ct
            f.argtypes = [P, P, ct]
        L.oracle_xform_matrix.restype = None
        L.oracle_xform_matrix.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.oracle_preround_f32.restype = ctypes.c_float
        L.oracle_preround_f32.argtypes = [ctypes.c_float, ctypes.c_float]
        _lib = L
    return _lib


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def _ptr(a):
    ct = ctypes.c_float if a.dtype == np.float32 else ctypes.c_double
    return a.ctypes.data_as(ctypes.POINTER(ct))


def _ptr_array(arrays, dtype):
    ct = ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double
    P = ctypes.POINTER(ct)
    return (P * len(arrays))(*[a.ctypes.data_as(P) for a in arrays])


# ---------------------------------------------------------------- primitives
def dct8(v, inverse=False, dtype=np.float32):
    """8-point scaled DCT-II (forward) / DCT-III (inverse) over the last axis."""
    a = np.ascontiguousarray(np.asarray(v, dtype=dtype).reshape(-1, 8)).copy()
    getattr(lib(), f"oracle_dct8_{_sfx(dtype)}")(_ptr(a), a.shape[0], int(inverse))
    return a.reshape(np.shape(v))


# Option codes of the NVRTC variant (the C ABI's BM3D_BM_* / BM3D_TRANSFORM_*).
BM_ERRORS = {"ssd": 0, "sad": 1, "zssd": 2, "zsad": 3, "ssd/norm": 4}
TRANSFORMS = {"dct": 0, "haar": 1, "wht": 2, "bior1.5": 3}


def transform3d(g, inverse=False, dtype=np.float32, t2=0, t1=0):
    """3-D transform of one 8x8x8 group [member, row, column]: transform kind
    t2 along rows and columns, t1 along the members."""
    a = np.ascontiguousarray(np.asarray(g, dtype=dtype)).copy()
    assert a.shape == (8, 8, 8)
    if t2 == 0 and t1 == 0:
        getattr(lib(), f"oracle_transform3d_{_sfx(dtype)}")(_ptr(a), int(inverse))
    else:
        getattr(lib(), f"oracle_transform3d_kinds_{_sfx(dtype)}")(_ptr(a), int(inverse), int(t2), int(t1))
    return a


def xform8(v, kind, inverse=False, dtype=np.float32):
    """8-point transform `kind` (TRANSFORMS) over the last axis."""
    a = np.ascontiguousarray(np.asarray(v, dtype=dtype).reshape(-1, 8)).copy()
    getattr(lib(), f"oracle_xform8_{_sfx(dtype)}")(_ptr(a), a.shape[0], int(kind), int(inverse))
    return a.reshape(np.shape(v))


def xform_matrix(kind, inverse=False):
    """The oracle's 8x8 matrix of transform `kind` (float64)."""
    out = np.zeros(64, np.float64)
    lib().oracle_xform_matrix(int(kind), int(inverse), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    return out.reshape(8, 8)


def _at(im, x, y, dtype):
    P = ctypes.POINTER(ctypes.c_float if np.dtype(dtype) == np.float32 else ctypes.c_double)
    return ctypes.cast(im.ctypes.data + (y * im.shape[1] + x) * im.itemsize, P)


def ssd(ref_block, img, x, y, dtype=np.float32):
    r = np.ascontiguousarray(np.asarray(ref_block, dtype=dtype).reshape(8, 8))
    im = np.ascontiguousarray(np.asarray(img, dtype=dtype))
    return getattr(lib(), f"oracle_ssd_{_sfx(dtype)}")(_ptr(r), _at(im, x, y, dtype), im.shape[1])


def block_cost(ref_block, img, x, y, cost, dtype=np.float32):
    """Block distance `cost` (BM_ERRORS) between ref_block and img[y:y+8, x:x+8]."""
    r = np.ascontiguousarray(np.asarray(ref_block, dtype=dtype).reshape(8, 8))
    im = np.ascontiguousarray(np.asarray(img, dtype=dtype))
    return getattr(lib(), f"oracle_block_cost_{_sfx(dtype)}")(_ptr(r), _at(im, x, y, dtype), im.shape[1],
                                                                int(cost))


def collab_hard(group, sigma_scaled, dtype=np.float32):
    """Hard-threshold collaborative filtering of one group [8, 8, 8]:
    returns (estimate, weight)."""
    g = np.ascontiguousarray(np.asarray(group, dtype=dtype)).copy()
    w = getattr(lib(), f"oracle_collab_hard_{_sfx(dtype)}")(_ptr(g), float(sigma_scaled))
    return g, w


def collab_wiener(group, basic, sigma_scaled, dtype=np.float32):
    """Wiener collaborative filtering of one group with its basic estimate."""
    g = np.ascontiguousarray(np.asarray(group, dtype=dtype)).copy()
    b = np.ascontiguousarray(np.asarray(basic, dtype=dtype)).copy()
    w = getattr(lib(), f"oracle_collab_wiener_{_sfx(dtype)}")(_ptr(g), _ptr(b), float(sigma_scaled))
    return g, w


def preround(v, E):
    return lib().oracle_preround_f32(float(v), float(E))


# ------------------------------------------------------------ parameters
def scale_sigma(sigma, final):
    """source/source.cpp:520-523, evaluated in float left to right."""
    f = np.float32(3.0) / np.float32(4.0) / np.float32(255.0) * np.float32(64.0) * \
        (np.float32(1.0) if final else np.float32(2.7))
    return np.float32(np.float32(sigma) * f)


@dataclass
class Params:
    sigma: tuple = (3.0, 3.0, 3.0)
    block_step: int = 8
    bm_range: int = 9
    radius: int = 0
    ps_num: int = 2
    ps_range: int = 4
    chroma: bool = False
    extractor_exp: int = 0
    bm_error: int = 0       # BM_ERRORS
    transform_2d: int = 0   # TRANSFORMS
    transform_1d: int = 0

    @property
    def extractor(self):
        return float(np.ldexp(np.float32(1.0), self.extractor_exp)) if self.extractor_exp else 0.0


def block_positions(n, bs):
    """cpu_source/source.cpp:787-791: min(k*bs, n-8) for k*bs < n-8+bs."""
    return [min(v, n - 8) for v in range(0, n - 8 + bs, bs)]


# --------------------------------------------------------------- matching
def match(slots, bs, bm_range, radius=0, ps_num=2, ps_range=4, dtype=np.float32, cost=0):
    """Group lists [blocks, 8, 3] (x, y, z) and errors [blocks, 8] for a window
    of 2r+1 planes (slots[r] is the centre). Also returns the candidate count."""
    slots = [np.ascontiguousarray(s, dtype=dtype) for s in slots]
    H, W = slots[0].shape
    nb = len(block_positions(H, bs)) * len(block_positions(W, bs))
    groups = np.zeros((nb, 8, 3), np.int32)
    errs = np.zeros((nb, 8), dtype)
    n = getattr(lib(), f"oracle_match_ex_{_sfx(dtype)}")(
        _ptr_array(slots, dtype), W, H, W, bs, bm_range, radius, ps_num, ps_range,
        groups.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), _ptr(errs), int(cost))
    return groups, errs, n


# ------------------------------------------------------------ BM3D(n)
def bm3d_window(src_slots, ref_slots, sigma_scaled, bs, bm_range, radius, ps_num, ps_range,
                chroma, final, extractor=0.0, dtype=np.float32, want_groups=False, options=(0, 0, 0)):
    """One kernel evaluation. src_slots/ref_slots: arrays [planes, 2r+1, H, W].
    Returns res [planes, 2r+1, 2, H, W] (num, den) and optionally groups."""
    src_slots = np.ascontiguousarray(src_slots, dtype=dtype)
    npl, tw, H, W = src_slots.shape
    assert tw == 2 * radius + 1 and npl == (3 if chroma else 1)
    res = np.zeros((npl, tw, 2, H, W), dtype)
    src_list = [src_slots[p, z] for p in range(npl) for z in range(tw)]
    if final:
        ref_slots = np.ascontiguousarray(ref_slots, dtype=dtype)
        ref_list = [ref_slots[p, z] for p in range(npl) for z in range(tw)]
    else:
        ref_list = src_list
    sig = (ctypes.c_float * 3)(*[float(s) for s in (list(sigma_scaled) + [0.0] * 3)[:3]])
    nb = len(block_positions(H, bs)) * len(block_positions(W, bs))
    groups = np.zeros((nb, 8, 3), np.int32) if want_groups else None
    gptr = groups.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) if want_groups else None
    n = getattr(lib(), f"oracle_bm3d_ex_{_sfx(dtype)}")(
        _ptr(res), _ptr_array(src_list, dtype), _ptr_array(ref_list, dtype),
        W, H, W, sig, bs, bm_range, radius, ps_num, ps_range, int(chroma), int(final),
        float(extractor), gptr, *[int(o) for o in options])
    if want_groups:
        return res, groups, n
    return res


def _window(clip, n, radius):
    N = clip.shape[0]
    idx = [min(max(n - radius + z, 0), N - 1) for z in range(2 * radius + 1)]
    return np.stack([clip[i] for i in idx], axis=1)  # [planes, 2r+1, H, W]


def bm3d_frame(clip, ref, n, p: Params, dtype=np.float32):
    """BM3D output frame n (source/source.cpp:206-441) for a clip
    [frames, planes, H, W]. radius 0: the denoised planes (unprocessed planes
    copied from clip). radius > 0: the stacked frame [planes, 2r+1, 2, H, W]
    (unprocessed planes zero, as with zero_init)."""
    final = ref is not None
    N, npl, H, W = clip.shape
    sig = [float(s) for s in (list(p.sigma) + [p.sigma[-1]] * 3)[:3]]
    process = [s >= FLT_EPSILON for s in sig]
    scaled = [scale_sigma(s, final) for s in sig]
    r = p.radius
    opts = (p.bm_error, p.transform_2d, p.transform_1d)
    if p.chroma:
        src_w = _window(clip, n, r)
        ref_w = _window(ref, n, r) if final else None
        res = bm3d_window(src_w, ref_w, scaled, p.block_step, p.bm_range, r, p.ps_num,
                          p.ps_range, True, final, p.extractor, dtype, options=opts)
        res_all = res
    else:
        res_all = np.zeros((npl, 2 * r + 1, 2, H, W), dtype)
        for pl in range(npl):
            if not process[pl]:
                continue
            src_w = _window(clip[:, pl:pl + 1], n, r)
            ref_w = _window(ref[:, pl:pl + 1], n, r) if final else None
            res_all[pl:pl + 1] = bm3d_window(src_w, ref_w, [scaled[pl]], p.block_step,
                                             p.bm_range, r, p.ps_num, p.ps_range, False, final,
                                             p.extractor, dtype, options=opts)
    if r == 0:
        out = np.array(clip[n], dtype=dtype, copy=True)
        for pl in range(npl):
            if process[pl]:
                out[pl] = res_all[pl, 0, 0] / res_all[pl, 0, 1]
        return out
    for pl in range(npl):
        if not process[pl]:
            res_all[pl] = 0
    return res_all


def vaggregate(stacks, src_frame, n, N, radius, process, dtype=np.float32):
    """VAggregateGetFrame (source/source.cpp:761-858). stacks: list of the 2r+1
    BM3D outputs for frames clamp(n - r + i) [planes, 2r+1, 2, H, W]. Slot
    s(i) = clamp(2r - i, n - N + 1 + r, n + r); sums in ascending i from 0."""
    out = np.array(src_frame, dtype=dtype, copy=True)
    for pl in range(out.shape[0]):
        if not process[pl]:
            continue
        num = np.zeros(out.shape[1:], dtype)
        den = np.zeros(out.shape[1:], dtype)
        for i in range(2 * radius + 1):
            s = min(max(2 * radius - i, n - N + 1 + radius), n + radius)
            num = num + stacks[i][pl, s, 0]
            den = den + stacks[i][pl, s, 1]
        out[pl] = num / den
    return out


def bm3dv2(clip, ref=None, p: Params = Params(), dtype=np.float32):
    """BM3Dv2 (source/source.cpp:911-989) over a whole clip [frames, planes, H, W]:
    BM3D, then VAggregate over the processed planes when radius > 0."""
    N, npl = clip.shape[:2]
    sig = [float(s) for s in (list(p.sigma) + [p.sigma[-1]] * 3)[:3]]
    process = [s >= FLT_EPSILON for s in sig][:npl]
    if not any(process):
        return np.array(clip, dtype=dtype, copy=True)
    if p.radius == 0:
        return np.stack([bm3d_frame(clip, ref, n, p, dtype) for n in range(N)])
    stacks = [bm3d_frame(clip, ref, n, p, dtype) for n in range(N)]
    r = p.radius
    out = []
    for n in range(N):
        win = [stacks[min(max(n - r + i, 0), N - 1)] for i in range(2 * r + 1)]
        out.append(vaggregate(win, clip[n], n, N, r, process, dtype))
    return np.stack(out)


def two_pass(clip, p: Params, dtype=np.float32):
    basic = bm3dv2(clip, None, p, dtype)
    final = bm3dv2(clip, basic.astype(dtype), p, dtype)
    return basic, final
