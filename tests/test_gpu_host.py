"""Host-frame route (bm3d_filter_get_frames_host): frames in host memory, as
a VapourSynth plugin holds them, staged through pinned buffers, pipelined in
chunks, on up to 4 resource sets. Bar: bit-identical to the device route
(bm3d_filter_get_frames) on the same frames -- the kernels are the same;
extractor_exp is set so that aggregation order cannot matter -- and, for the
two-pass BM3Dv2, bit-identical to the oracle's two-pass pipeline."""
import ctypes
import threading

import numpy as np
import pytest

from bm3d_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _vs():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import vs
    return vs


KW = dict(sigma=[10.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=10)


@pytest.mark.parametrize("N", [5, 40])
def test_bm3dv2_two_pass_host_equals_device(N):
    """N = 40 runs in three pipelined chunks."""
    vs = _vs()
    _, noisy = synth.make_clip(40, 56, N, 1, noise_sd=10.0, seed=90, motion=True)
    dev = vs.Clip.from_array(noisy, "GRAY")
    host = vs.Clip.from_array(noisy, "GRAY", device="cpu")
    b_d = vs.BM3Dv2(dev, **KW)
    f_d = vs.BM3Dv2(dev, ref=b_d, **KW)
    b_h = vs.BM3Dv2(host, **KW)
    assert b_h.on_host
    f_h = vs.BM3Dv2(host, ref=b_h, **KW)
    np.testing.assert_array_equal(b_h.numpy(), b_d.numpy())
    np.testing.assert_array_equal(f_h.numpy(), f_d.numpy())
    # and both equal the oracle's two-pass BM3Dv2 (exact sums with extractor_exp)
    p = O.Params(sigma=tuple(KW["sigma"]), block_step=KW["block_step"], bm_range=KW["bm_range"],
                 radius=KW["radius"], ps_num=KW["ps_num"], ps_range=KW["ps_range"],
                 extractor_exp=KW["extractor_exp"])
    want_b, want_f = O.two_pass(noisy, p)
    np.testing.assert_array_equal(b_h.numpy(), want_b)
    np.testing.assert_array_equal(f_h.numpy(), want_f)


@pytest.mark.parametrize("radius", [0, 2])
def test_bm3d_stacks_and_vaggregate_host(radius):
    vs = _vs()
    _, noisy = synth.make_clip(32, 40, 6, 3, noise_sd=10.0, seed=91, motion=True, yuv=True)
    kw = dict(KW, sigma=[10.0, 0.0, 7.0], radius=radius)
    dev, host = vs.Clip.from_array(noisy, "YUV"), vs.Clip.from_array(noisy, "YUV", device="cpu")
    s_d, s_h = vs.BM3D(dev, **kw), vs.BM3D(host, **kw)
    np.testing.assert_array_equal(s_h.numpy(), s_d.numpy())
    if radius:
        assert s_h.props["BM3D_V_radius"] == radius
        v_d = vs.VAggregate(s_d, dev, [0, 2])
        v_h = vs.VAggregate(s_h, host, [0, 2])
        np.testing.assert_array_equal(v_h.numpy(), v_d.numpy())


def test_chroma_and_subsampled_host():
    vs = _vs()
    _, noisy = synth.make_clip(40, 48, 4, 3, noise_sd=10.0, seed=92, motion=True, yuv=True)
    kw = dict(KW, sigma=[10.0, 8.0, 12.0], chroma=True)
    d = vs.BM3Dv2(vs.Clip.from_array(noisy, "YUV"), **kw).numpy()
    h = vs.BM3Dv2(vs.Clip.from_array(noisy, "YUV", device="cpu"), **kw).numpy()
    np.testing.assert_array_equal(h, d)
    planes = [synth.make_clip(hh, ww, 4, 1, noise_sd=10, seed=93 + p)[1][:, 0]
              for p, (hh, ww) in enumerate([(48, 64), (24, 32), (24, 32)])]
    d = vs.BM3Dv2(vs.Clip.from_planes(planes, "YUV", (1, 1)), **KW)
    h = vs.BM3Dv2(vs.Clip.from_planes(planes, "YUV", (1, 1), device="cpu"), **KW)
    for p in range(3):
        np.testing.assert_array_equal(h.planes[p].numpy(), d.planes[p].cpu().numpy())


def _filter(vs, create, kw, clip, ref=None):
    from bm3d_amd import _capi as C
    a = vs._args(**{k: kw.get(k) for k in ("sigma", "block_step", "bm_range", "radius", "ps_num", "ps_range",
                                           "chroma", "device_id", "fast", "extractor_exp", "zero_init",
                                           "precision")})
    h = ctypes.c_void_p()
    rvi = ref.video_info() if ref is not None else None
    vs._check(create(ctypes.byref(a), ctypes.byref(clip.video_info()), ctypes.byref(rvi) if rvi else None,
                     ctypes.byref(h)))
    return h, C


def _host_frames(C, arrs, lo, hi):
    """bm3d_host_frame entries over per-plane numpy arrays [N, H, Wpadded]."""
    out = (C.HostFrame * (hi - lo + 1))()
    for i in range(hi - lo + 1):
        for p, a in enumerate(arrs):
            out[i].data[p] = a[lo + i].ctypes.data
            out[i].stride[p] = a.strides[1] // 4
    return out


@pytest.mark.parametrize("fast", [1, 0])
def test_per_frame_calls_from_threads_padded_rows(fast):
    """Per-frame calls as VapourSynth makes them, from 4 threads sharing one
    filter, on host rows padded to a wider pitch."""
    vs = _vs()
    N, H, W = 8, 40, 56
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=10.0, seed=94, motion=True)
    kw = dict(KW, fast=fast)
    want = vs.BM3Dv2(vs.Clip.from_array(noisy, "GRAY"), **KW).numpy()
    padded = np.zeros((N, H, W + 24), np.float32)
    padded[:, :, :W] = noisy[:, 0]
    outbuf = np.full((N, H, W + 8), np.nan, np.float32)
    clip = vs.Clip.from_array(noisy, "GRAY", device="cpu")
    h, C = _filter(vs, C_lib().bm3dv2_filter_create, kw, clip)
    errors = []

    def work(frames):
        for n in frames:
            lo, hi = ctypes.c_int(), ctypes.c_int()
            rc = C.lib().bm3d_filter_host_range(h, n, 1, ctypes.byref(lo), ctypes.byref(hi))
            fin = _host_frames(C, [padded], lo.value, hi.value)
            fout = _host_frames(C, [outbuf], n, n)
            rc = rc or C.lib().bm3d_filter_get_frames_host(h, fin, None, lo.value, fout, n, 1)
            if rc:
                errors.append((n, C.lib().bm3d_last_error().decode()))

    threads = [threading.Thread(target=work, args=(list(range(t, N, 4)),)) for t in range(4)]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    C.lib().bm3d_filter_destroy(h)
    assert not errors, errors
    np.testing.assert_array_equal(outbuf[:, :, :W], want[:, 0])
    assert np.isnan(outbuf[:, :, W:]).all()  # the pitch padding is not written


def C_lib():
    from bm3d_amd import _capi as C
    return C.lib()


def test_host_range_and_errors():
    vs = _vs()
    from bm3d_amd import _capi as C
    _, noisy = synth.make_clip(32, 40, 10, 1, noise_sd=10.0, seed=95)
    clip = vs.Clip.from_array(noisy, "GRAY", device="cpu")
    h, _ = _filter(vs, C.lib().bm3dv2_filter_create, dict(KW, radius=2), clip)
    lo, hi = ctypes.c_int(), ctypes.c_int()
    assert C.lib().bm3d_filter_host_range(h, 5, 2, ctypes.byref(lo), ctypes.byref(hi)) == 0
    assert (lo.value, hi.value) == (1, 9)  # BM3Dv2 reads 2r frames either side
    assert C.lib().bm3d_filter_host_range(h, 0, 1, ctypes.byref(lo), ctypes.byref(hi)) == 0
    assert (lo.value, hi.value) == (0, 4)
    fin = clip.host_frames(0, 4)
    out = vs.Clip.from_array(np.zeros_like(noisy), "GRAY", device="cpu").host_frames(0, 0)
    assert C.lib().bm3d_filter_get_frames_host(h, fin, None, 1, out, 0, 1) == C.BM3D_EINVAL
    assert "frame lo" in C.lib().bm3d_last_error().decode()
    fin[2].data[0] = None
    assert C.lib().bm3d_filter_get_frames_host(h, fin, None, 0, out, 0, 1) == C.BM3D_EINVAL
    assert "host frame 2 plane 0" in C.lib().bm3d_last_error().decode()
    C.lib().bm3d_filter_destroy(h)
