"""Plan-level Python wrapper over the C ABI, for device-resident torch tensors.

A clip tensor is float32 on a HIP device, shaped [frames, planes, height,
width] and contiguous (row stride = width). torch only provides the device
memory and the stream; all work runs in libbm3d_mi355x.so.
"""
from __future__ import annotations

import ctypes

import torch

from . import _capi as C


def _stream_ptr(stream):
    if stream is None:
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def _check_clip(t, name):
    if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or not t.is_cuda:
        raise TypeError(f"{name} must be a float32 HIP tensor")
    if t.dim() != 4 or not t.is_contiguous():
        raise ValueError(f"{name} must be a contiguous [frames, planes, height, width] tensor")


def clip_desc(t):
    n, p, h, w = t.shape
    return C.Clip(ctypes.c_void_p(t.data_ptr()), p * h * w, h * w, n)


# Options of the NVRTC variant (rtc_source/source.cpp:854-912): names as the
# plugin accepts them, codes as the C ABI's BM3D_BM_* / BM3D_TRANSFORM_*.
BM_ERRORS = {"ssd": 0, "sad": 1, "zssd": 2, "zsad": 3, "ssd/norm": 4}
TRANSFORMS = {"dct": 0, "haar": 1, "wht": 2, "bior1.5": 3}


def _option(table, value, name):
    if isinstance(value, str):
        key = value.lower()
        if key not in table:
            raise ValueError(f"invalid '{name}': {key}")
        return table[key]
    if int(value) not in table.values():
        raise ValueError(f"invalid '{name}': {value}")
    return int(value)


def _desc(width, height, sigma, block_step=8, bm_range=9, radius=0, ps_num=2, ps_range=4, chroma=False,
          final=False, extractor_exp=0, precision="fp32", stride=None, bm_error="ssd", transform_2d="dct",
          transform_1d="dct", mixed_temporal_keys=0):
    sig = list(sigma) if isinstance(sigma, (list, tuple)) else [sigma]
    sig = (sig + [sig[-1]] * 3)[:3]
    d = C.Desc()
    d.width, d.height = int(width), int(height)
    d.stride = int(stride if stride is not None else width)
    for i in range(3):
        d.sigma[i] = float(sig[i])
    d.block_step, d.bm_range, d.radius = int(block_step), int(bm_range), int(radius)
    d.ps_num, d.ps_range = int(ps_num), int(ps_range)
    d.chroma, d.final_, d.extractor_exp = int(bool(chroma)), int(bool(final)), int(extractor_exp)
    d.precision = {"fp32": C.PRECISION_FP32, "mixed": C.PRECISION_MIXED}[precision]
    d.bm_error = _option(BM_ERRORS, bm_error, "bm_error_s")
    d.transform_2d = _option(TRANSFORMS, transform_2d, "transform_2d_s")
    d.transform_1d = _option(TRANSFORMS, transform_1d, "transform_1d_s")
    d.mixed_temporal_keys = int(mixed_temporal_keys)
    return d


class Plan:
    """One plane group's baked parameters (the role of get_graphexec,
    source/kernel.cu:759-864)."""

    def __init__(self, width, height, sigma, block_step=8, bm_range=9, radius=0, ps_num=2,
                 ps_range=4, chroma=False, final=False, extractor_exp=0, precision="fp32",
                 stride=None, bm_error="ssd", transform_2d="dct", transform_1d="dct", mixed_temporal_keys=0):
        d = _desc(width, height, sigma, block_step, bm_range, radius, ps_num, ps_range, chroma, final,
                  extractor_exp, precision, stride, bm_error, transform_2d, transform_1d, mixed_temporal_keys)
        self.desc = d
        self.nplanes = 3 if chroma else 1
        self.radius = int(radius)
        self.final = bool(final)
        h = ctypes.c_void_p()
        C.check(C.lib().bm3d_plan_create(ctypes.byref(d), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            C.lib().bm3d_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stack_size(self):
        return int(C.lib().bm3d_stack_size(self._h))

    @property
    def num_blocks(self):
        return int(C.lib().bm3d_num_blocks(self._h))

    def new_stacks(self, count, device=None):
        d = self.desc
        return torch.empty((count, self.nplanes, 2 * self.radius + 1, 2, d.height, d.stride),
                           dtype=torch.float32, device=device or "cuda")

    def execute(self, src, ref=None, first=0, count=None, out=None, stream=None):
        """Stacks [count, planes, 2r+1, 2, H, W] for output frames first.. of src."""
        _check_clip(src, "clip")
        if self.final:
            if ref is None:
                raise ValueError("final pass needs ref")
            _check_clip(ref, "ref")
        n = src.shape[0]
        count = n - first if count is None else count
        if out is None:
            out = self.new_stacks(count, src.device)
        s = clip_desc(src)
        r = clip_desc(ref) if ref is not None else None
        C.check(C.lib().bm3d_execute(self._h, ctypes.c_void_p(out.data_ptr()), 0, 0,
                                     ctypes.byref(s), ctypes.byref(r) if r else None,
                                     int(first), int(count), _stream_ptr(stream)))
        return out

    def execute_vagg(self, src, ref=None, first=0, count=None, out=None, stream=None):
        """BM3D followed by VAggregate (radius > 0) for frames first ..
        first + count - 1, without materialising the stacks:
        [count, planes, H, W] float32."""
        _check_clip(src, "clip")
        n = src.shape[0]
        count = n - first if count is None else count
        d = self.desc
        if out is None:
            out = torch.empty((count, self.nplanes, d.height, d.width), dtype=torch.float32, device=src.device)
        s = clip_desc(src)
        r = clip_desc(ref) if ref is not None else None
        C.check(C.lib().bm3d_execute_vagg(self._h, ctypes.c_void_p(out.data_ptr()), out.stride(0),
                                          out.stride(1), out.stride(2), ctypes.byref(s),
                                          ctypes.byref(r) if r else None, int(first), int(count),
                                          _stream_ptr(stream)))
        return out

    def uses_mixed_kernel(self):
        """True when block matching runs on the mixed-precision kernel."""
        return bool(C.lib().bm3d_plan_matching_kernel(self._h))

    def uses_matrix_core_search(self):
        """True when the mixed kernel's spatial search may use the matrix cores."""
        return C.lib().bm3d_plan_matching_kernel(self._h) == 2

    def uses_strip_search(self):
        """True when the mixed kernel's spatial search runs by strips shared
        between lanes (block steps 2 and 4; the default there)."""
        return C.lib().bm3d_plan_matching_kernel(self._h) == 3

    def uses_mixed_filter(self):
        """True when collaborative filtering runs on the mixed-precision kernel."""
        return bool(C.lib().bm3d_plan_filter_kernel(self._h))

    def select_kernels(self, matching, filter):
        """Selects the matching and filter kernels: each "fp32" or "mixed";
        matching "mixed" is the mixed kernel with a per-lane spatial search,
        "mixed-mfma" with its spatial search's cross terms on the matrix
        cores, "mixed-strip" with its spatial search by strips shared between
        lanes (where the block step allows; the default)."""
        codes = {"fp32": 0, "mixed": 1}
        mcodes = dict(codes, **{"mixed-mfma": 2, "mixed-strip": 3})
        if matching not in mcodes or filter not in codes:
            raise ValueError("kernel must be 'fp32' or 'mixed' (matching also 'mixed-mfma' or 'mixed-strip')")
        C.check(C.lib().bm3d_plan_select_kernels(self._h, mcodes[matching], codes[filter]))

    @property
    def tuning(self):
        """Launch shapes as a dict (match_rows, filter_tile_x, filter_tile_y,
        filter_waves, batch_frames, chunks, match_width)."""
        t = C.Tuning()
        C.check(C.lib().bm3d_plan_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        """Changes some launch shapes; the others keep their values."""
        t = C.Tuning()
        C.check(C.lib().bm3d_plan_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        C.check(C.lib().bm3d_plan_set_tuning(self._h, ctypes.byref(t)))
        return self.tuning

    def autotune(self, src, ref=None, first=0, count=None, stream=None):
        """Times candidate launch shapes on these frames and keeps the fastest."""
        _check_clip(src, "clip")
        n = src.shape[0]
        count = n - first if count is None else count
        s = clip_desc(src)
        r = clip_desc(ref) if ref is not None else None
        C.check(C.lib().bm3d_plan_autotune(self._h, ctypes.byref(s), ctypes.byref(r) if r else None,
                                           int(first), int(count), _stream_ptr(stream)))
        return self.tuning

    def count_candidates(self, src, ref=None, first=0, count=None, stream=None):
        """Exact number of in-plane block-matching candidates for the frames."""
        _check_clip(src, "clip")
        n = src.shape[0]
        count = n - first if count is None else count
        s = clip_desc(src)
        r = clip_desc(ref) if ref is not None else None
        out = ctypes.c_ulonglong()
        C.check(C.lib().bm3d_count_candidates(self._h, ctypes.byref(s), ctypes.byref(r) if r else None,
                                              int(first), int(count), ctypes.byref(out),
                                              _stream_ptr(stream)))
        return int(out.value)

    def match(self, src, ref=None, first=0, count=None, stream=None):
        """Packed group lists [count, nblocks, 8] (int32 view of x | y<<14 | z<<27)."""
        _check_clip(src, "clip")
        n = src.shape[0]
        count = n - first if count is None else count
        g = torch.empty((count, self.num_blocks, 8), dtype=torch.int32, device=src.device)
        s = clip_desc(src)
        r = clip_desc(ref) if ref is not None else None
        C.check(C.lib().bm3d_match(self._h, ctypes.c_void_p(g.data_ptr()), ctypes.byref(s),
                                   ctypes.byref(r) if r else None, int(first), int(count),
                                   _stream_ptr(stream)))
        return g


class Batch:
    """Many clips denoised together (bm3d_batch_*): each job is a dict with
    the Plan parameters (width, height, sigma, block_step, bm_range, radius,
    ps_num, ps_range, chroma, extractor_exp, precision, bm_error,
    transform_2d, transform_1d, mixed_temporal_keys), `src` (a clip
    tensor), `passes` (1: BM3Dv2; 2: two-pass BM3Dv2) and optional `out` /
    `basic` tensors shaped like src (allocated when absent). execute()
    replays the batch's captured graph."""

    def __init__(self, jobs, scratch_budget=0):
        arr = (C.BatchJob * len(jobs))()
        self.jobs = []
        for i, j in enumerate(jobs):
            src = j["src"]
            _check_clip(src, "clip")
            n, planes, h, w = src.shape
            passes = int(j.get("passes", 2))
            out = j.get("out")
            if out is None:
                out = torch.empty_like(src)
            basic = j.get("basic")
            if basic is None and passes == 2:
                basic = torch.empty_like(src)
            plan_kw = {k: j[k] for k in ("sigma", "block_step", "bm_range", "radius", "ps_num", "ps_range",
                                         "chroma", "extractor_exp", "precision", "bm_error", "transform_2d",
                                         "transform_1d", "mixed_temporal_keys") if k in j}
            d = _desc(w, h, **plan_kw)
            arr[i].desc = d
            arr[i].passes = passes
            arr[i].src = clip_desc(src)
            arr[i].basic = ctypes.c_void_p(basic.data_ptr() if basic is not None else 0)
            arr[i].out = ctypes.c_void_p(out.data_ptr())
            self.jobs.append(dict(src=src, out=out, basic=basic, passes=passes))
        h = ctypes.c_void_p()
        C.check(C.lib().bm3d_batch_create(arr, len(jobs), int(scratch_budget), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            C.lib().bm3d_batch_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, stream=None):
        C.check(C.lib().bm3d_batch_execute(self._h, _stream_ptr(stream)))
        return [j["out"] for j in self.jobs]

    @property
    def tuning(self):
        t = C.Tuning()
        C.check(C.lib().bm3d_batch_get_tuning(self._h, ctypes.byref(t)))
        return t.as_dict()

    def set_tuning(self, **kw):
        t = C.Tuning()
        C.check(C.lib().bm3d_batch_get_tuning(self._h, ctypes.byref(t)))
        for k, v in kw.items():
            if not hasattr(t, k):
                raise KeyError(k)
            setattr(t, k, int(v))
        C.check(C.lib().bm3d_batch_set_tuning(self._h, ctypes.byref(t)))
        return self.tuning

    def autotune(self, stream=None):
        C.check(C.lib().bm3d_batch_autotune(self._h, _stream_ptr(stream)))
        return self.tuning

    @property
    def info(self):
        waves, launches, scratch = ctypes.c_int(), ctypes.c_int(), ctypes.c_longlong()
        C.check(C.lib().bm3d_batch_info(self._h, ctypes.byref(waves), ctypes.byref(launches),
                                        ctypes.byref(scratch)))
        return {"waves": waves.value, "launches": launches.value, "scratch_bytes": scratch.value}


def unpack_groups(g):
    """[.., 8] packed -> [.., 8, 3] (x, y, z) int64 numpy."""
    import numpy as np
    a = np.asarray(g.cpu() if isinstance(g, torch.Tensor) else g).astype(np.int64) & 0xFFFFFFFF
    x = a & ((1 << 14) - 1)
    y = (a >> 14) & ((1 << 13) - 1)
    z = a >> 27
    return np.stack([x, y, z], axis=-1)


def aggregate(stacks, out=None, stream=None):
    """r = 0: out[f, p] = num / den for stacks [count, planes, 1, 2, H, W]."""
    cnt, npl, tw, two, h, w = stacks.shape
    assert tw == 1 and two == 2
    if out is None:
        out = torch.empty((cnt, npl, h, w), dtype=torch.float32, device=stacks.device)
    C.check(C.lib().bm3d_aggregate(ctypes.c_void_p(out.data_ptr()), npl * h * w, h * w, w,
                                   ctypes.c_void_p(stacks.data_ptr()), 0, 0, w, h, w, npl, cnt,
                                   _stream_ptr(stream)))
    return out


def vaggregate(stacks, radius, num_frames, first=0, count=None, stack_first=0, out=None,
               stream=None):
    """VAggregate over stacks [m, planes, 2r+1, 2, H, W] of frames stack_first..;
    returns [count, planes, H, W] for output frames first..first+count-1."""
    m, npl, tw, two, h, w = stacks.shape
    assert tw == 2 * radius + 1 and two == 2
    count = num_frames - first if count is None else count
    if out is None:
        out = torch.empty((count, npl, h, w), dtype=torch.float32, device=stacks.device)
    C.check(C.lib().bm3d_vaggregate(ctypes.c_void_p(out.data_ptr()), npl * h * w, h * w, w,
                                    ctypes.c_void_p(stacks.data_ptr()), 0, 0, int(stack_first),
                                    w, h, w, npl, int(radius), int(num_frames), int(first),
                                    int(count), _stream_ptr(stream)))
    return out
