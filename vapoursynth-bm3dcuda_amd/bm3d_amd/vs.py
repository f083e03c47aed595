"""The plugin surface of bm3dcuda -- BM3D, VAggregate, BM3Dv2 -- over device
clips, with VapourSynth-style keyword arguments.

    from bm3d_amd import vs
    src = vs.Clip.from_array(frames, "YUV")          # [N, 3, H, W] float32
    basic = vs.BM3Dv2(src, sigma=[10, 10, 10], radius=1)
    final = vs.BM3Dv2(src, ref=basic, sigma=[10, 10, 10], radius=1)

Argument handling (per-plane arrays, defaults, checks, error messages) and
frame semantics are those of the reference plugin (source/source.cpp:
457-1027); they are implemented in C++ (csrc/bm3d_filters.cpp) behind the
C ABI, and this module only marshals clips and arguments.
"""
from __future__ import annotations

import ctypes

import numpy as np
import torch

from . import _capi as C
from .core import _stream_ptr

FAMILIES = {"GRAY": 1, "RGB": 2, "YUV": 3}


class Error(C.BM3DError):
    """Raised where VapourSynth would report a filter error."""


class Clip:
    """Float32 planes on a HIP device, or in host memory (device="cpu": the
    filters then take the host-frame route, bm3d_filter_get_frames_host, as
    a VapourSynth plugin whose frames live in host memory would). planes[p]
    is [frames, H_p, W_p]; when all planes share one size they are views of
    one [frames, P, H, W] tensor. props holds the clip-level frame
    properties the filters set."""

    def __init__(self, planes, family="GRAY", subsampling=(0, 0), props=None, base=None):
        self.planes = list(planes)
        self.family = family
        self.subsampling = tuple(subsampling)
        self.props = dict(props or {})
        self.base = base

    @classmethod
    def from_array(cls, arr, family="GRAY", device="cuda"):
        t = torch.as_tensor(np.asarray(arr, dtype=np.float32) if not isinstance(arr, torch.Tensor) else arr)
        t = t.to(device=device, dtype=torch.float32).contiguous()
        if t.dim() == 3:
            t = t.unsqueeze(1)
        return cls([t[:, p] for p in range(t.shape[1])], family, (0, 0), base=t)

    @classmethod
    def from_planes(cls, planes, family="YUV", subsampling=(1, 1), device="cuda"):
        ps = [torch.as_tensor(np.asarray(p, dtype=np.float32)).to(device).contiguous() for p in planes]
        return cls(ps, family, subsampling)

    @property
    def num_frames(self):
        return self.planes[0].shape[0]

    @property
    def height(self):
        return self.planes[0].shape[1]

    @property
    def width(self):
        return self.planes[0].shape[2]

    def numpy(self):
        if self.base is not None:
            return self.base.cpu().numpy()
        return [p.cpu().numpy() for p in self.planes]

    def video_info(self):
        vi = C.VideoInfo()
        vi.color_family = FAMILIES[self.family]
        vi.sample_type = 1
        vi.bits_per_sample = 32
        vi.subsampling_w, vi.subsampling_h = self.subsampling
        vi.num_planes = len(self.planes)
        vi.width, vi.height = self.width, self.height
        vi.num_frames = self.num_frames
        return vi

    @property
    def on_host(self):
        return self.planes[0].device.type == "cpu"

    def frames(self):
        fr = C.Frames()
        for p, t in enumerate(self.planes[:3]):
            fr.data[p] = t.data_ptr()
            fr.frame_stride[p] = t.stride(0)
            fr.stride[p] = t.stride(1)
        return fr

    def host_frames(self, lo, hi):
        """bm3d_host_frame entries of frames lo .. hi (host clips)."""
        arr = (C.HostFrame * (hi - lo + 1))()
        for i in range(hi - lo + 1):
            for p, t in enumerate(self.planes[:3]):
                arr[i].data[p] = t[lo + i].data_ptr()
                arr[i].stride[p] = t.stride(1)
        return arr


def _alloc_like(vi, family, subsampling, device):
    n, h, w = vi.num_frames, vi.height, vi.width
    if subsampling == (0, 0):
        base = torch.empty((n, vi.num_planes, h, w), dtype=torch.float32, device=device)
        return Clip([base[:, p] for p in range(vi.num_planes)], family, subsampling, base=base)
    planes = []
    for p in range(vi.num_planes):
        ph = h if p == 0 else h >> subsampling[1]
        pw = w if p == 0 else w >> subsampling[0]
        planes.append(torch.empty((n, ph, pw), dtype=torch.float32, device=device))
    return Clip(planes, family, subsampling)


def _args(sigma, block_step, bm_range, radius, ps_num, ps_range, chroma, device_id, fast,
          extractor_exp, zero_init, precision, bm_error_s=None, transform_2d_s=None, transform_1d_s=None):
    a = C.Args()

    def arr(name, vals, conv):
        if vals is None:
            return
        vals = list(vals) if isinstance(vals, (list, tuple)) else [vals]
        setattr(a, "n_" + name, len(vals))
        buf = getattr(a, name)
        for i, v in enumerate(vals[:3]):
            buf[i] = conv(v)

    def one(name, val, conv=int):
        if val is None:
            return
        setattr(a, "n_" + name, 1)
        setattr(a, name, conv(val))

    arr("sigma", sigma, float)
    arr("block_step", block_step, int)
    arr("bm_range", bm_range, int)
    one("radius", radius)
    arr("ps_num", ps_num, int)
    arr("ps_range", ps_range, int)
    one("chroma", chroma)
    one("device_id", device_id)
    one("fast", fast)
    one("extractor_exp", extractor_exp)
    one("zero_init", zero_init)
    if precision is not None:
        one("precision", {"fp32": C.PRECISION_FP32, "mixed": C.PRECISION_MIXED}.get(precision, precision))
    for name, vals in (("bm_error_s", bm_error_s), ("transform_2d_s", transform_2d_s),
                       ("transform_1d_s", transform_1d_s)):
        arr(name, vals, lambda v: v.encode() if isinstance(v, str) else bytes(v))
    return a


def _check(rc):
    if rc != C.BM3D_OK:
        raise Error(rc, C.lib().bm3d_last_error().decode())


def _run(h, clip, second, stream):
    vi = C.VideoInfo()
    radius = ctypes.c_int()
    process = (ctypes.c_int * 3)()
    _check(C.lib().bm3d_filter_info(h, ctypes.byref(vi), ctypes.byref(radius), process))
    out = _alloc_like(vi, clip.family, clip.subsampling, clip.planes[0].device)
    if clip.on_host:
        # the VapourSynth route: host frames in, host frames out
        if second is not None and not second.on_host:
            raise ValueError("clip and ref/src must both be host clips or both device clips")
        lo, hi = ctypes.c_int(), ctypes.c_int()
        n = clip.num_frames
        _check(C.lib().bm3d_filter_host_range(h, 0, n, ctypes.byref(lo), ctypes.byref(hi)))
        f_in = clip.host_frames(lo.value, hi.value)
        f_sec = second.host_frames(lo.value, hi.value) if second is not None else None
        f_out = out.host_frames(0, n - 1)
        _check(C.lib().bm3d_filter_get_frames_host(h, f_in, f_sec, lo.value, f_out, 0, n))
        return out, radius.value, list(process)
    f_in = clip.frames()
    f_sec = second.frames() if second is not None else None
    f_out = out.frames()
    _check(C.lib().bm3d_filter_get_frames(h, ctypes.byref(f_in), ctypes.byref(f_sec) if f_sec else None,
                                          ctypes.byref(f_out), 0, clip.num_frames, _stream_ptr(stream)))
    return out, radius.value, list(process)


def _filter(create, clip, ref, kw, stream):
    a = _args(**kw)
    h = ctypes.c_void_p()
    rvi = ref.video_info() if ref is not None else None
    cvi = clip.video_info()
    _check(create(ctypes.byref(a), ctypes.byref(cvi), ctypes.byref(rvi) if rvi else None, ctypes.byref(h)))
    try:
        return _run(h, clip, ref, stream)
    finally:
        C.lib().bm3d_filter_destroy(h)


def BM3D(clip, ref=None, sigma=None, block_step=None, bm_range=None, radius=None, ps_num=None,
         ps_range=None, chroma=None, device_id=None, fast=None, extractor_exp=None, zero_init=None,
         precision=None, stream=None, bm_error_s=None, transform_2d_s=None, transform_1d_s=None):
    """bm3dcuda.BM3D (source/source.cpp:457-720). With radius > 0 the output
    frames are the stacked weighted sums and weights (height 2(2r+1)H) with
    props BM3D_V_radius / BM3D_V_process, for VAggregate. bm_error_s,
    transform_2d_s and transform_1d_s are the bm3dcuda_rtc options
    (rtc_source/source.cpp:854-912)."""
    kw = dict(sigma=sigma, block_step=block_step, bm_range=bm_range, radius=radius, ps_num=ps_num,
              ps_range=ps_range, chroma=chroma, device_id=device_id, fast=fast,
              extractor_exp=extractor_exp, zero_init=zero_init, precision=precision,
              bm_error_s=bm_error_s, transform_2d_s=transform_2d_s, transform_1d_s=transform_1d_s)
    out, r, process = _filter(C.lib().bm3d_filter_create, clip, ref, kw, stream)
    if r:
        out.props.update(BM3D_V_radius=r, BM3D_V_process=process)
    return out


def VAggregate(clip, src, planes, stream=None):
    """bm3dcuda.VAggregate (source/source.cpp:722-909)."""
    planes = list(planes) if isinstance(planes, (list, tuple)) else [planes]
    pl = (ctypes.c_longlong * max(1, len(planes)))(*planes)
    h = ctypes.c_void_p()
    cvi, svi = clip.video_info(), src.video_info()
    _check(C.lib().bm3d_vaggregate_create(pl, len(planes), ctypes.byref(cvi), ctypes.byref(svi),
                                          ctypes.byref(h)))
    try:
        out, _, _ = _run(h, clip, src, stream)
        return out
    finally:
        C.lib().bm3d_filter_destroy(h)


def BM3Dv2(clip, ref=None, sigma=None, block_step=None, bm_range=None, radius=None, ps_num=None,
           ps_range=None, chroma=None, device_id=None, fast=None, extractor_exp=None,
           zero_init=None, precision=None, stream=None, bm_error_s=None, transform_2d_s=None,
           transform_1d_s=None):
    """bm3dcuda.BM3Dv2 (source/source.cpp:911-989): BM3D then VAggregate;
    with the bm3dcuda_rtc options as BM3D."""
    kw = dict(sigma=sigma, block_step=block_step, bm_range=bm_range, radius=radius, ps_num=ps_num,
              ps_range=ps_range, chroma=chroma, device_id=device_id, fast=fast,
              extractor_exp=extractor_exp, zero_init=zero_init, precision=precision,
              bm_error_s=bm_error_s, transform_2d_s=transform_2d_s, transform_1d_s=transform_1d_s)
    out, _, _ = _filter(C.lib().bm3dv2_filter_create, clip, ref, kw, stream)
    return out
