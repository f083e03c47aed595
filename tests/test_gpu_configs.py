"""The workload configurations beside the benchmark, at their real sizes.

* 640x480 GRAYS basic BM3D, radius 0, block_step 8, sigma 10, no
  extractor: the whole frame against the CPU oracle (FP32), and the mixed
  path within 0.05 dB of the FP64 oracle.
* YUV444 1920x1080, 5 frames, chroma, two-pass BM3D (radius 0), sigma 10,
  block_step 4, extractor_exp 6: the full frames on the GPU against the
  oracle on a 256x256 crop whose origin is a multiple of block_step. In the
  crop's interior (64 pixels from its edges) every reference block, search
  window and group member -- and, for the final pass, every basic-estimate
  pixel they read -- is the same as in the full frame, so with extractor_exp
  the values agree bit for bit. Two runs are bitwise identical.
* YUV444 3840x2160, 16 frames, chroma, two-pass BM3Dv2 radius 2,
  block_step 4, bm_range 16, ps_num 2, ps_range 6, extractor_exp 8, noise
  20/255: the mixed path at full size is bitwise run-to-run deterministic
  and within 0.05 dB of the FP32 path per plane, and its group lists equal
  the FP32 path's except among exactly tied candidates; on a 5-frame 128x128 crop
  clip it is within 0.05 dB of the FP64 oracle per plane (max-abs stated).
"""
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


def _psnr_planes(a, clean):
    return [synth.psnr(a[:, p], clean[:, p]) for p in range(a.shape[1])]


def test_config_640x480_basic_r0():
    vs = _vs()
    clean, noisy = synth.make_clip(480, 640, 1, 1, noise_sd=10.0, seed=0)
    src = vs.Clip.from_array(noisy, "GRAY")
    kw = dict(sigma=10.0, block_step=8, bm_range=9)
    got = vs.BM3D(src, **kw).numpy()
    prm = O.Params(sigma=(10.0,), block_step=8, bm_range=9)
    want = O.bm3dv2(noisy, None, prm)
    np.testing.assert_allclose(got, want, rtol=0, atol=2e-5)
    f64 = O.bm3dv2(noisy.astype(np.float64), None, prm, dtype=np.float64)
    mixed = vs.BM3D(src, precision="mixed", **kw).numpy()
    d = synth.psnr(mixed, clean) - synth.psnr(f64, clean)
    assert abs(d) <= 0.05, d
    assert synth.psnr(got, clean) > synth.psnr(noisy, clean) + 3.0


def test_config_yuv444_1080p_chroma_two_pass():
    vs = _vs()
    H, W, N = 1080, 1920, 5
    clean, noisy = synth.make_clip(H, W, N, 3, noise_sd=10.0, seed=1, motion=False, yuv=True)
    src = vs.Clip.from_array(noisy, "YUV")
    kw = dict(sigma=[10.0, 10.0, 10.0], block_step=4, bm_range=9, chroma=True, extractor_exp=6)
    basic = vs.BM3Dv2(src, **kw)
    final = vs.BM3Dv2(src, ref=basic, **kw)
    gb, gf = basic.numpy(), final.numpy()
    # run-to-run bitwise determinism with extractor_exp
    basic2 = vs.BM3Dv2(src, **kw)
    final2 = vs.BM3Dv2(src, ref=basic2, **kw)
    np.testing.assert_array_equal(final2.numpy(), gf)
    # crop comparison against the oracle (frames 0 and 4; radius 0 frames are independent)
    y0, x0, S, m = 400, 800, 256, 64
    prm = O.Params(sigma=(10.0, 10.0, 10.0), block_step=4, bm_range=9, chroma=True, extractor_exp=6)
    for n in (0, N - 1):
        crop = np.ascontiguousarray(noisy[n:n + 1, :, y0:y0 + S, x0:x0 + S])
        ob, of = O.two_pass(crop, prm)
        np.testing.assert_array_equal(gb[n:n + 1, :, y0 + m:y0 + S - m, x0 + m:x0 + S - m],
                                      ob[:, :, m:S - m, m:S - m], err_msg=f"basic frame {n}")
        np.testing.assert_array_equal(gf[n:n + 1, :, y0 + m:y0 + S - m, x0 + m:x0 + S - m],
                                      of[:, :, m:S - m, m:S - m], err_msg=f"final frame {n}")
    for p, (a, b) in enumerate(zip(_psnr_planes(gf, clean), _psnr_planes(noisy, clean))):
        assert a > b + 3.0, (p, a, b)


def test_config_4k_yuv444_r2_chroma():
    vs = _vs()
    H, W, N = 2160, 3840, 16
    clean, noisy = synth.make_clip(H, W, N, 3, noise_sd=20.0, seed=4, motion=True, yuv=True)
    src = vs.Clip.from_array(noisy, "YUV")
    kw = dict(sigma=[20.0, 20.0, 20.0], block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=6,
              chroma=True, extractor_exp=8)

    def two_pass(prec, clip):
        b = vs.BM3Dv2(clip, precision=prec, **kw)
        return vs.BM3Dv2(clip, ref=b, precision=prec, **kw).numpy()

    mixed = two_pass("mixed", src)
    np.testing.assert_array_equal(two_pass("mixed", src), mixed)  # bitwise deterministic
    # basic-pass group lists of three frames: the mixed matcher's equal the
    # FP32 matcher's except among exactly tied candidates
    from bm3d_amd import core
    import mixed_ties
    mk = dict(block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=6, chroma=True)
    plan_m = core.Plan(W, H, [20.0] * 3, precision="mixed", **mk)
    plan_32 = core.Plan(W, H, [20.0] * 3, **mk)
    src_t = torch.from_numpy(noisy).cuda()
    for n in (0, 7, 15):
        gm = core.unpack_groups(plan_m.match(src_t, first=n, count=1))
        g32 = core.unpack_groups(plan_32.match(src_t, first=n, count=1))
        nd, nl = mixed_ties.tie_only(gm, g32, noisy, 2, 4, first=n)
        print(f"frame {n}: {nd} of {nl} lists differ, all at exact ties")
    del plan_m, plan_32, src_t
    fp32 = two_pass("fp32", src)
    pm, pf = _psnr_planes(mixed, clean), _psnr_planes(fp32, clean)
    pn = _psnr_planes(noisy, clean)
    for p in range(3):
        assert abs(pm[p] - pf[p]) <= 0.05, (p, pm[p], pf[p])
        assert pm[p] > pn[p] + 3.0, (p, pm[p], pn[p])
    del mixed, fp32, src
    # per-plane error against the FP64 oracle on a crop clip
    y0, x0, S, F = 1000, 1800, 128, 5
    crop_n = np.ascontiguousarray(noisy[:F, :, y0:y0 + S, x0:x0 + S])
    crop_c = np.ascontiguousarray(clean[:F, :, y0:y0 + S, x0:x0 + S])
    prm = O.Params(sigma=(20.0, 20.0, 20.0), block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=6,
                   chroma=True, extractor_exp=8)
    _, f64 = O.two_pass(crop_n.astype(np.float64), prm, dtype=np.float64)
    got = two_pass("mixed", vs.Clip.from_array(crop_n, "YUV"))
    for p in range(3):
        d = synth.psnr(got[:, p], crop_c[:, p]) - synth.psnr(f64[:, p], crop_c[:, p])
        err = float(np.abs(got[:, p] - f64[:, p]).max())
        print(f"plane {p}: delta PSNR vs FP64 oracle {d:+.5f} dB, max-abs {err:.5f}")
        assert abs(d) <= 0.05, (p, d)
