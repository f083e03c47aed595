"""Plugin-level filters (bm3d_amd.vs: BM3D, VAggregate, BM3Dv2) on the GPU
against the oracle's restatement of the reference plugin's frame semantics
(source/source.cpp:206-441, 746-862, 911-989). extractor_exp is set where
bit-identity is asserted (the pre-rounded sums are exact)."""
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


def _yuv(H, W, N, seed, sd=10.0):
    clean, noisy = synth.make_clip(H, W, N, 3, noise_sd=sd, seed=seed, motion=True, yuv=True)
    return clean, noisy


def test_bm3dv2_two_pass_per_plane_params_bitexact():
    vs = _vs()
    H, W, N = 40, 48, 4
    _, noisy = _yuv(H, W, N, 1)
    src = vs.Clip.from_array(noisy, "YUV")
    kw = dict(sigma=[10.0, 6.0, 8.0], block_step=[4, 8], bm_range=[4, 5, 3], radius=1, ps_num=[2, 3],
              ps_range=[2], extractor_exp=10)
    basic = vs.BM3Dv2(src, **kw)
    final = vs.BM3Dv2(src, ref=basic, **kw)
    got_b, got_f = basic.numpy(), final.numpy()
    # tail-filled per-plane parameters, each plane its own clip in the oracle
    per_plane = [(4, 4, 2, 2), (8, 5, 3, 2), (8, 3, 3, 2)]
    for p, (bs, R, psn, psr) in enumerate(per_plane):
        prm = O.Params(sigma=(kw["sigma"][p],), block_step=bs, bm_range=R, radius=1, ps_num=psn,
                       ps_range=psr, extractor_exp=10)
        b, f = O.two_pass(np.ascontiguousarray(noisy[:, p:p + 1]), prm)
        np.testing.assert_array_equal(got_b[:, p:p + 1], b)
        np.testing.assert_array_equal(got_f[:, p:p + 1], f)


def test_bm3dv2_chroma_bitexact():
    vs = _vs()
    H, W, N = 40, 48, 3
    _, noisy = _yuv(H, W, N, 2)
    src = vs.Clip.from_array(noisy, "YUV")
    kw = dict(sigma=[10.0, 8.0, 12.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2,
              chroma=True, extractor_exp=10)
    basic = vs.BM3Dv2(src, **kw)
    final = vs.BM3Dv2(src, ref=basic, **kw)
    prm = O.Params(sigma=(10.0, 8.0, 12.0), block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2,
                   chroma=True, extractor_exp=10)
    b, f = O.two_pass(noisy, prm)
    np.testing.assert_array_equal(basic.numpy(), b)
    np.testing.assert_array_equal(final.numpy(), f)


def test_bm3d_r0_unprocessed_planes_copied():
    vs = _vs()
    _, noisy = _yuv(32, 40, 2, 3)
    src = vs.Clip.from_array(noisy, "YUV")
    out = vs.BM3D(src, sigma=[10.0, 0.0], block_step=4, bm_range=4, extractor_exp=10).numpy()
    np.testing.assert_array_equal(out[:, 1:], noisy[:, 1:])
    want = O.bm3dv2(np.ascontiguousarray(noisy[:, :1]), None,
                    O.Params(sigma=(10.0,), block_step=4, bm_range=4, extractor_exp=10))
    np.testing.assert_array_equal(out[:, :1], want)


def test_bm3d_radius_stacks_props_and_vaggregate():
    vs = _vs()
    H, W, N, r = 32, 40, 4, 1
    _, noisy = _yuv(H, W, N, 4)
    src = vs.Clip.from_array(noisy, "YUV")
    kw = dict(sigma=[10.0, 0.0, 7.0], block_step=4, bm_range=4, radius=r, ps_num=2, ps_range=2,
              extractor_exp=10)
    st = vs.BM3D(src, **kw)
    assert st.height == H * 2 * (2 * r + 1)
    assert st.props == {"BM3D_V_radius": r, "BM3D_V_process": [1, 0, 1]}
    stacks = st.numpy()
    assert not stacks[:, 1].any()  # zero_init
    agg = vs.VAggregate(st, src, planes=[0, 2]).numpy()
    v2 = vs.BM3Dv2(src, **kw).numpy()
    np.testing.assert_array_equal(agg, v2)
    np.testing.assert_array_equal(agg[:, 1], noisy[:, 1])
    # the stack of one frame equals the oracle's window evaluation
    n = 2
    win = np.stack([noisy[min(max(n - r + z, 0), N - 1), 0:1] for z in range(2 * r + 1)], axis=1)
    want = O.bm3d_window(win, None, [O.scale_sigma(10.0, False)], 4, 4, r, 2, 2, False, False, 1024.0)
    np.testing.assert_array_equal(stacks[n, 0].reshape(2 * r + 1, 2, H, W), want[0])


def test_bm3dv2_subsampled_yuv420():
    vs = _vs()
    H, W, N = 48, 64, 3
    planes_clean = [synth.make_clip(h, w, N, 1, noise_sd=10, seed=5 + p)[1][:, 0]
                    for p, (h, w) in enumerate([(H, W), (H // 2, W // 2), (H // 2, W // 2)])]
    src = vs.Clip.from_planes(planes_clean, "YUV", (1, 1))
    kw = dict(sigma=[10.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=10)
    out = vs.BM3Dv2(src, **kw)
    for p in range(3):
        prm = O.Params(sigma=(10.0,), block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2,
                       extractor_exp=10)
        want = O.bm3dv2(planes_clean[p][:, None], None, prm)[:, 0]
        np.testing.assert_array_equal(out.planes[p].cpu().numpy(), want)


def test_bm3dv2_nothing_processed_is_passthrough():
    vs = _vs()
    _, noisy = _yuv(32, 40, 2, 6)
    src = vs.Clip.from_array(noisy, "YUV")
    out = vs.BM3Dv2(src, sigma=[0.0], radius=1).numpy()
    np.testing.assert_array_equal(out, noisy)


def test_bm3dv2_mixed_precision_matches_fp32():
    vs = _vs()
    clean, noisy = synth.make_clip(64, 80, 4, 1, noise_sd=15, seed=7, motion=True)
    src = vs.Clip.from_array(noisy, "GRAY")
    kw = dict(sigma=15.0, block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3, extractor_exp=10)
    a = vs.BM3Dv2(src, **kw).numpy()
    b = vs.BM3Dv2(src, precision="mixed", **kw).numpy()
    # FP16 pre-scoring with exact rescoring, FP16 transforms with FP32
    # accumulation: close to the FP32 output, and as close to the clean clip
    d = (a - b).astype(np.float64)
    assert np.sqrt(np.mean(d ** 2)) < 2e-3
    assert np.abs(d).max() < 0.05
    assert abs(synth.psnr(a, clean) - synth.psnr(b, clean)) < 0.05


def test_runtime_errors_raise():
    vs = _vs()
    _, noisy = synth.make_clip(32, 40, 2, 1, noise_sd=10, seed=8)
    src = vs.Clip.from_array(noisy, "GRAY")
    with pytest.raises(vs.Error, match='"ps_num" must be in range'):
        vs.BM3D(src, ps_num=9)
    with pytest.raises(vs.Error, match="YUV444"):
        vs.BM3D(src, chroma=True)
