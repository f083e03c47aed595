"""GPU parity of the NVRTC variant's options (bm_error_s, transform_2d_s,
transform_1d_s; rtc_source/kernel.hpp:103-504, rtc_source/source.cpp:854-912)
against the CPU oracle.

Bars (as for the default options):
* group lists: identical for every block distance (the FP32 matcher
  evaluates each distance in the oracle's operation order);
* stacks with extractor_exp set: bit-identical for every transform pair;
* stacks without extractor_exp: rtol 2e-6 + atol 5e-6 (fixed-point and
  float-atomic accumulation order);
* mixed precision (FP16 MFMA transforms): output PSNR within 0.05 dB of the
  FP64 oracle's for the invertible transforms. The reference's Haar inverse
  is not an inverse (it reads coefficients 2 and 4 in place of 3, 5, 6, 7),
  so the Haar output is far from the clean image for every implementation;
  there the bar is the mixed output's distance to the FP64 oracle output.
"""
import numpy as np
import pytest

from bm3d_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _core():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import core
    return core


def _window(clip, n, r):
    N = clip.shape[0]
    return np.stack([clip[min(max(n - r + z, 0), N - 1)] for z in range(2 * r + 1)], axis=1)


def _oracle_stacks(noisy, ref, prm, final, chroma, opts, dtype=np.float32):
    N = noisy.shape[0]
    r = prm["radius"]
    sig = [O.scale_sigma(s, final) for s in prm["sigma"]]
    E = float(np.ldexp(np.float32(1), prm["extractor_exp"])) if prm["extractor_exp"] else 0.0
    out, groups = [], []
    for n in range(N):
        res, g, _ = O.bm3d_window(_window(noisy, n, r), _window(ref, n, r) if final else None, sig,
                                  prm["block_step"], prm["bm_range"], r, prm["ps_num"], prm["ps_range"], chroma,
                                  final, E, dtype=dtype, want_groups=True, options=opts)
        out.append(res)
        groups.append(g)
    return np.stack(out), np.stack(groups)


def _plan(core, W, H, prm, final, chroma, opts, precision="fp32"):
    return core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], prm["radius"], prm["ps_num"],
                     prm["ps_range"], chroma=chroma, final=final, extractor_exp=prm["extractor_exp"],
                     precision=precision, bm_error=opts[0], transform_2d=opts[1], transform_1d=opts[2])


CASES = {
    "r1": (40, 56, 4, 1, False, dict(sigma=[15.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=3,
                                      extractor_exp=0)),
    "r0-bs3": (45, 61, 1, 1, False, dict(sigma=[15.0], block_step=3, bm_range=4, radius=0, ps_num=2, ps_range=4,
                                          extractor_exp=0)),
    "chroma-r1": (40, 48, 3, 3, True, dict(sigma=[10.0, 8.0, 12.0], block_step=4, bm_range=4, radius=1,
                                           ps_num=2, ps_range=2, extractor_exp=0)),
}

# (bm_error, transform_2d, transform_1d)
OPTIONS = [(1, 1, 1), (2, 2, 2), (3, 3, 3), (4, 0, 3), (0, 3, 0), (0, 0, 1), (2, 1, 2), (4, 2, 0)]


@pytest.mark.parametrize("cost", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("name", list(CASES))
def test_match_groups_identical(name, cost):
    core = _core()
    H, W, N, P, chroma, prm = CASES[name]
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=40 + cost, motion=True, yuv=chroma)
    plan = _plan(core, W, H, prm, False, chroma, (cost, 0, 0))
    got = core.unpack_groups(plan.match(torch.from_numpy(noisy).cuda()))
    torch.cuda.synchronize()
    _, want = _oracle_stacks(noisy, None, prm, False, chroma, (cost, 0, 0))
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("extractor_exp", [0, 12])
@pytest.mark.parametrize("opts", OPTIONS, ids=[f"e{a}-t{b}{c}" for a, b, c in OPTIONS])
@pytest.mark.parametrize("name", list(CASES))
def test_two_pass_stacks(name, opts, extractor_exp):
    core = _core()
    H, W, N, P, chroma, prm = CASES[name]
    prm = dict(prm, extractor_exp=extractor_exp)
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=50, motion=True, yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    got_b = _plan(core, W, H, prm, False, chroma, opts).execute(src).cpu().numpy()
    want_b, _ = _oracle_stacks(noisy, None, prm, False, chroma, opts)
    if extractor_exp:
        np.testing.assert_array_equal(got_b, want_b)
    else:
        np.testing.assert_allclose(got_b, want_b, rtol=2e-6, atol=5e-6)
    # final pass with a shared basic estimate: the oracle's, through VAggregate
    r = prm["radius"]
    basic = np.empty_like(noisy)
    keep = [p for p in range(P) if not (chroma and prm["sigma"][p] == 0.0)]
    for n in range(N):
        win = [want_b[min(max(n - r + i, 0), N - 1)] for i in range(2 * r + 1)]
        basic[n] = O.vaggregate(win, noisy[n], n, N, r, [p in keep for p in range(P)]) if r else \
            np.where(want_b[n, :, 0, 1] > 0, want_b[n, :, 0, 0] / np.maximum(want_b[n, :, 0, 1], 1e-30), noisy[n])
    got_f = _plan(core, W, H, prm, True, chroma, opts).execute(src, torch.from_numpy(basic).cuda()).cpu().numpy()
    want_f, _ = _oracle_stacks(noisy, basic, prm, True, chroma, opts)
    if extractor_exp:
        np.testing.assert_array_equal(got_f, want_f)
    else:
        np.testing.assert_allclose(got_f, want_f, rtol=2e-6, atol=5e-6)


@pytest.mark.parametrize("opts", [(0, 2, 2), (0, 3, 3), (1, 3, 2), (0, 0, 3), (0, 1, 1)],
                         ids=["wht", "bior", "sad-bior-wht", "dct-bior", "haar"])
def test_mixed_precision_vs_fp64(opts):
    core = _core()
    H, W, N = 64, 80, 5
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=20.0, seed=60, motion=True)
    prm = dict(sigma=[20.0], block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3, extractor_exp=0)
    src = torch.from_numpy(noisy).cuda()
    plan_b = _plan(core, W, H, prm, False, False, opts, "mixed")
    plan_f = _plan(core, W, H, prm, True, False, opts, "mixed")
    assert plan_b.uses_mixed_filter() and plan_b.uses_mixed_kernel() == (opts[0] == 0)
    basic = plan_b.execute_vagg(src)
    final = plan_f.execute_vagg(src, basic).cpu().numpy()
    p = O.Params(sigma=(20.0,), block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3,
                 bm_error=opts[0], transform_2d=opts[1], transform_1d=opts[2])
    _, want = O.two_pass(noisy.astype(np.float64), p, np.float64)
    if opts[1] != 1 and opts[2] != 1:
        d = abs(synth.psnr(final, clean) - synth.psnr(want, clean))
        assert d <= 0.05, d
        assert synth.psnr(final, clean) > synth.psnr(noisy, clean) + 3.0
    else:
        assert synth.psnr(final, want) > 35.0


def test_mixed_matcher_serves_ssd_only():
    core = _core()
    prm = CASES["r1"][5]
    for cost in range(5):
        plan = _plan(core, 56, 40, prm, False, False, (cost, 0, 0), "mixed")
        assert plan.uses_mixed_kernel() == (cost == 0)
    plan = _plan(core, 56, 40, prm, False, False, (2, 0, 0), "mixed")
    from bm3d_amd import _capi as C
    with pytest.raises(C.BM3DError) as ei:
        plan.select_kernels("mixed", "mixed")
    assert ei.value.code == C.BM3D_EUNSUPPORTED and "ssd" in str(ei.value)


def test_fused_path_and_batch_with_options():
    """bm3d_execute_vagg and the batch launcher give the stacks path's values
    for non-default options (extractor_exp set: bit for bit)."""
    core = _core()
    H, W, N = 40, 56, 5
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=70, motion=True)
    src = torch.from_numpy(noisy).cuda()
    prm = dict(sigma=[15.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=3, extractor_exp=11)
    jobs = []
    wants = []
    for opts in [(3, 2, 3), (4, 1, 0), (0, 3, 3)]:
        plan = _plan(core, W, H, prm, False, False, opts)
        want = core.vaggregate(plan.execute(src), 1, N).cpu().numpy()
        np.testing.assert_array_equal(plan.execute_vagg(src).cpu().numpy(), want)
        wants.append(want)
        jobs.append(dict(src=src, passes=1, sigma=[15.0], block_step=4, bm_range=4, radius=1, ps_num=2,
                         ps_range=3, extractor_exp=11, bm_error=opts[0], transform_2d=opts[1],
                         transform_1d=opts[2]))
    batch = core.Batch(jobs)
    outs = batch.execute()
    torch.cuda.synchronize()
    for o, w in zip(outs, wants):
        np.testing.assert_array_equal(o.cpu().numpy(), w)


def test_plugin_filters_with_string_options():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import vs
    H, W, N = 40, 48, 4
    _, noisy = synth.make_clip(H, W, N, 3, noise_sd=10.0, seed=80, motion=True, yuv=True)
    src = vs.Clip.from_array(noisy, "YUV")
    kw = dict(sigma=[10.0, 6.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=10,
              bm_error_s=["ZSSD", "sad"], transform_2d_s=["wht", "Bior1.5"], transform_1d_s=["haar", "dct", "wht"])
    basic = vs.BM3Dv2(src, **kw)
    final = vs.BM3Dv2(src, ref=basic, **kw)
    # tail fill: plane 2 takes plane 1's bm_error_s and transform_2d_s
    per_plane = [(2, 2, 1), (1, 3, 0), (1, 3, 2)]
    for p, (e, t2, t1) in enumerate(per_plane):
        prm = O.Params(sigma=((10.0, 6.0, 6.0)[p],), block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2,
                       extractor_exp=10, bm_error=e, transform_2d=t2, transform_1d=t1)
        b, f = O.two_pass(np.ascontiguousarray(noisy[:, p:p + 1]), prm)
        np.testing.assert_array_equal(basic.numpy()[:, p:p + 1], b)
        np.testing.assert_array_equal(final.numpy()[:, p:p + 1], f)
    with pytest.raises(vs.Error, match="invalid 'transform_1d_s': dst"):
        vs.BM3D(src, transform_1d_s="DST")


def test_mixed_batch_with_options_equals_plans():
    """Mixed-precision jobs with different transform pairs share one filter
    launch of a batch (their MFMA operands travel with each job): bit-equal
    to each plan's fused path with extractor_exp."""
    core = _core()
    H, W, N = 40, 56, 5
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=71, motion=True)
    src = torch.from_numpy(noisy).cuda()
    prm = dict(sigma=[15.0], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=3, extractor_exp=11)
    jobs, wants = [], []
    for opts in [(0, 0, 0), (0, 3, 2), (2, 2, 3), (0, 1, 0)]:
        pb = _plan(core, W, H, prm, False, False, opts, "mixed")
        pf = _plan(core, W, H, prm, True, False, opts, "mixed")
        b = pb.execute_vagg(src)
        wants.append(pf.execute_vagg(src, b).cpu().numpy())
        jobs.append(dict(src=src, passes=2, precision="mixed", sigma=[15.0], block_step=4, bm_range=4, radius=1,
                         ps_num=2, ps_range=3, extractor_exp=11, bm_error=opts[0], transform_2d=opts[1],
                         transform_1d=opts[2]))
    batch = core.Batch(jobs)
    outs = batch.execute()
    torch.cuda.synchronize()
    for o, w in zip(outs, wants):
        np.testing.assert_array_equal(o.cpu().numpy(), w)
