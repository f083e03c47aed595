"""Argument handling of the plugin-level filters (bm3d_filter_create,
bm3dv2_filter_create, bm3d_vaggregate_create) against the checks and
messages of the reference's BM3DCreate / VAggregateCreate
(source/source.cpp:476-611, 880-909). All checks before the device check
run without a GPU."""
import ctypes

import pytest

from bm3d_amd import _capi as C


def vi(family=1, sample_type=1, bits=32, ssw=0, ssh=0, planes=1, w=64, h=48, n=5):
    v = C.VideoInfo()
    v.color_family, v.sample_type, v.bits_per_sample = family, sample_type, bits
    v.subsampling_w, v.subsampling_h, v.num_planes = ssw, ssh, planes
    v.width, v.height, v.num_frames = w, h, n
    return v


def args(**kw):
    a = C.Args()
    for k, vals in kw.items():
        if isinstance(vals, (list, tuple)):
            setattr(a, "n_" + k, len(vals))
            buf = getattr(a, k)
            for i, x in enumerate(vals):
                buf[i] = x
        else:
            setattr(a, "n_" + k, 1)
            setattr(a, k, vals)
    return a


def create(a, clip, ref=None, v2=False):
    h = ctypes.c_void_p()
    fn = C.lib().bm3dv2_filter_create if v2 else C.lib().bm3d_filter_create
    rc = fn(ctypes.byref(a), ctypes.byref(clip), ctypes.byref(ref) if ref else None, ctypes.byref(h))
    msg = C.lib().bm3d_last_error().decode()
    if h.value:
        C.lib().bm3d_filter_destroy(h)
    return rc, msg


CASES = [
    (dict(), vi(sample_type=0, bits=8), None, "BM3D: only constant format 32bit float input supported"),
    (dict(), vi(bits=16), None, "BM3D: only constant format 32bit float input supported"),
    (dict(), vi(w=0), None, "BM3D: only constant format 32bit float input supported"),
    (dict(), vi(), vi(family=3, planes=3), 'BM3D: "ref" must be of the same format as "clip"'),
    (dict(), vi(), vi(w=32), 'BM3D: "ref" must be of the same dimensions as "clip"'),
    (dict(), vi(), vi(n=4), 'BM3D: "ref" must be of the same number of frames as "clip"'),
    (dict(sigma=[3.0, -1.0]), vi(), None, 'BM3D: "sigma" must be non-negative'),
    (dict(block_step=[4, 0]), vi(), None, 'BM3D: "block_step" must be in range [1, 8]'),
    (dict(block_step=[9]), vi(), None, 'BM3D: "block_step" must be in range [1, 8]'),
    (dict(bm_range=[0]), vi(), None, 'BM3D: "bm_range" must be positive'),
    (dict(radius=-1), vi(), None, 'BM3D: "radius" must be non-negative'),
    (dict(ps_num=[2, 9]), vi(), None, 'BM3D: "ps_num" must be in range [1, 8]'),
    (dict(ps_range=[0]), vi(), None, 'BM3D: "ps_range" must be positive'),
    (dict(chroma=1), vi(), None, 'BM3D: clip format must be YUV444 when "chroma" is true'),
    (dict(chroma=1), vi(family=3, planes=3, ssw=1, ssh=1), None,
     'BM3D: clip format must be YUV444 when "chroma" is true'),
    (dict(device_id=-1), vi(), None, "BM3D: invalid device ID (-1)"),
    (dict(device_id=64), vi(), None, "BM3D: invalid device ID (64)"),
]


@pytest.mark.parametrize("v2", [False, True])
@pytest.mark.parametrize("kw,clip,ref,msg", CASES)
def test_create_rejects(kw, clip, ref, msg, v2):
    rc, got = create(args(**kw), clip, ref, v2)
    assert rc == C.BM3D_EINVAL
    assert got == msg


def test_check_order_matches_reference():
    # several bad arguments: the first check in the reference's order wins
    rc, msg = create(args(sigma=[-1.0], block_step=[0], bm_range=[0]), vi())
    assert msg == 'BM3D: "sigma" must be non-negative'
    rc, msg = create(args(block_step=[0], bm_range=[0], radius=-1), vi())
    assert msg == 'BM3D: "block_step" must be in range [1, 8]'


def test_v2_all_planes_unprocessed_needs_no_device():
    # BM3Dv2 returns the clip itself before BM3D is created (source/source.cpp:932-942)
    h = ctypes.c_void_p()
    a = args(sigma=[0.0])
    c = vi(family=3, planes=3)
    assert C.lib().bm3dv2_filter_create(ctypes.byref(a), ctypes.byref(c), None, ctypes.byref(h)) == 0
    out = C.VideoInfo()
    r = ctypes.c_int()
    proc = (ctypes.c_int * 3)()
    assert C.lib().bm3d_filter_info(h, ctypes.byref(out), ctypes.byref(r), proc) == 0
    assert (out.width, out.height, r.value, list(proc)) == (64, 48, 0, [0, 0, 0])
    C.lib().bm3d_filter_destroy(h)


def vagg(planes, clip, src):
    h = ctypes.c_void_p()
    pl = (ctypes.c_longlong * max(1, len(planes)))(*planes)
    rc = C.lib().bm3d_vaggregate_create(pl, len(planes), ctypes.byref(clip), ctypes.byref(src), ctypes.byref(h))
    msg = C.lib().bm3d_last_error().decode()
    return rc, msg, h


def test_vaggregate_radius_from_height():
    # radius = (H_clip / H_src - 2) / 4  (source/source.cpp:892)
    for r in (1, 2, 3):
        rc, msg, h = vagg([0], vi(h=48 * 2 * (2 * r + 1)), vi())
        assert rc == 0, msg
        rad = ctypes.c_int()
        out = C.VideoInfo()
        proc = (ctypes.c_int * 3)()
        C.lib().bm3d_filter_info(h, ctypes.byref(out), ctypes.byref(rad), proc)
        assert rad.value == r and out.height == 48 and list(proc) == [1, 0, 0]
        C.lib().bm3d_filter_destroy(h)


def test_vaggregate_rejects():
    rc, msg, _ = vagg([0], vi(h=48), vi())
    assert rc == C.BM3D_EINVAL and "radius > 0" in msg
    rc, msg, _ = vagg([3], vi(h=48 * 6), vi())
    assert rc == C.BM3D_EINVAL and msg == "VAggregate: invalid plane index"


# ---- the bm3dcuda_rtc string options (rtc_source/source.cpp:854-912) ----
RTC_CASES = [
    (dict(bm_error_s=[b"xyz"]), "BM3D: invalid 'bm_error_s': xyz"),
    (dict(bm_error_s=[b"SSD", b"Foo"]), "BM3D: invalid 'bm_error_s': foo"),
    (dict(bm_error_s=[b"ssd_norm"]), "BM3D: invalid 'bm_error_s': ssd_norm"),
    (dict(bm_error_s=[b""]), "BM3D: invalid 'bm_error_s': "),
    (dict(transform_2d_s=[b"dct", b"BIOR1.5", b"fft"]), "BM3D: invalid 'transform_2d_s': fft"),
    (dict(transform_2d_s=[b"bior1_5"]), "BM3D: invalid 'transform_2d_s': bior1_5"),
    (dict(transform_1d_s=[b"Haar", b"dst"]), "BM3D: invalid 'transform_1d_s': dst"),
    # bm_error_s is checked first, then transform_2d_s, then transform_1d_s
    (dict(bm_error_s=[b"x"], transform_2d_s=[b"y"], transform_1d_s=[b"z"]), "BM3D: invalid 'bm_error_s': x"),
    (dict(transform_2d_s=[b"y"], transform_1d_s=[b"z"]), "BM3D: invalid 'transform_2d_s': y"),
]


@pytest.mark.parametrize("v2", [False, True])
@pytest.mark.parametrize("kw,msg", RTC_CASES)
def test_rtc_options_rejected(kw, msg, v2):
    rc, got = create(args(**kw), vi(), None, v2)
    assert rc == C.BM3D_EINVAL
    assert got == msg


@pytest.mark.parametrize("kw", [
    dict(bm_error_s=[b"SSD/Norm"], transform_2d_s=[b"WHT"], transform_1d_s=[b"Bior1.5"]),
    dict(bm_error_s=[b"zsad", b"zssd", b"sad"], transform_2d_s=[b"haar", b"dct"]),
])
def test_rtc_options_accepted_case_insensitive(kw):
    # valid options pass on to the next check, the device (no GPU in the CPU
    # suite; with a GPU the filter is created)
    rc, got = create(args(device_id=64, **kw), vi())
    assert got == "BM3D: invalid device ID (64)"


def test_rtc_options_after_chroma_check():
    rc, got = create(args(chroma=1, bm_error_s=[b"bad"]), vi())
    assert got == 'BM3D: clip format must be YUV444 when "chroma" is true'


def test_v2_passthrough_host_frames_need_no_device():
    """BM3Dv2 with nothing to process returns its source (source/source.cpp:
    932-942); on host frames that is a host copy, so it runs without a GPU."""
    import numpy as np
    from bm3d_amd import vs
    rng = np.random.default_rng(1)
    arr = rng.uniform(size=(5, 3, 24, 40)).astype(np.float32)
    src = vs.Clip.from_array(arr, "YUV", device="cpu")
    out = vs.BM3Dv2(src, sigma=[0.0], radius=1)
    assert out.on_host
    np.testing.assert_array_equal(out.numpy(), arr)
    lo, hi = ctypes.c_int(), ctypes.c_int()
    h = ctypes.c_void_p()
    a = args(sigma=[0.0], radius=2)
    c = vi(family=3, planes=3, w=40, h=24)
    assert C.lib().bm3dv2_filter_create(ctypes.byref(a), ctypes.byref(c), None, ctypes.byref(h)) == 0
    assert C.lib().bm3d_filter_host_range(h, 2, 2, ctypes.byref(lo), ctypes.byref(hi)) == 0
    assert (lo.value, hi.value) == (2, 3)  # nothing processed: no temporal reach
    assert C.lib().bm3d_filter_host_range(h, 4, 2, ctypes.byref(lo), ctypes.byref(hi)) == C.BM3D_EINVAL
    C.lib().bm3d_filter_destroy(h)
