"""Search reaches beyond what the LDS tiles hold.

The reference accepts any positive bm_range and radius
(source/source.cpp:537-547). Here a plan's tiles shrink as the reach
M = bm_range + radius * ps_range grows (160 KB LDS per work-group):

* filter tiles fall back to 4 x 4 blocks with an accumulation margin below
  M; members beyond it are added with global float atomics straight into
  the stacks, and bm3d_execute_vagg refuses the plan (BM3Dv2 then runs the
  stacks and VAggregate);
* when even 4 rows of blocks' union search window exceed the LDS, the
  matcher reads the planes from global memory.

Geometry of the cases (bs = block_step; the filter tile and margin are the
largest that fit, as the plan's tiling chooses them):

  case          bs  bm_range  r  ps_range  M    matcher          filter margin
  r4-defaults    8      9     4     4      25   LDS, 16 rows     17
  r4-bs4         4      9     4     4      25   LDS, 16 rows     23
  R40-r1         8     40     1     4      44   LDS, 4 rows      42
  R40-r1-yuv     8     40     1     4      44   LDS, 4 rows      42  (chroma)
  R64-r0         8     64     0     4      64   plane reads      64
  R64-r1         8     64     1     4      68   plane reads      42
  R100-r0-zssd   8    100     0     4     100   plane reads      84  (zssd)

Bars, against the CPU oracle: group lists identical, in-plane candidate
counts equal, and stacks of both passes bit-identical with extractor_exp
(every pre-rounded sum is exact, so neither the atomic order nor the split
between LDS regions and global atomics can show); without extractor_exp the
tolerance of the other stack tests. The BM3Dv2 filter at radius 4 equals
the oracle's two-pass BM3Dv2 bit for bit, and its mixed-precision form is
within 0.05 dB of the FP64 oracle.
"""
import dataclasses

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


COSTS = ["ssd", "sad", "zssd", "zsad", "ssd/norm"]

# (H, W, N, planes, chroma, block_step, bm_range, radius, ps_num, ps_range, bm_error, region margin reduced)
CASES = {
    "r4-defaults": (48, 64, 10, 1, False, 8, 9, 4, 2, 4, "ssd", True),
    "r4-bs4": (48, 64, 9, 1, False, 4, 9, 4, 2, 4, "ssd", True),
    "R40-r1": (96, 112, 4, 1, False, 8, 40, 1, 2, 4, "ssd", True),
    "R40-r1-yuv": (96, 112, 3, 3, True, 8, 40, 1, 2, 4, "ssd", True),
    "R64-r0": (144, 160, 2, 1, False, 8, 64, 0, 2, 4, "ssd", False),
    "R64-r1": (144, 160, 3, 1, False, 8, 64, 1, 2, 4, "ssd", True),
    "R100-r0-zssd": (216, 232, 1, 1, False, 8, 100, 0, 2, 4, "zssd", True),
}


def _window(clip, n, r):
    N = clip.shape[0]
    return np.stack([clip[min(max(n - r + z, 0), N - 1)] for z in range(2 * r + 1)], axis=1)


def _oracle(noisy, ref, sig, case, final, E):
    H, W, N, P, chroma, bs, R, r, ps_num, ps_range, cost, _ = case
    out, groups, count = [], [], 0
    for n in range(N):
        res, g, c = O.bm3d_window(_window(noisy, n, r), _window(ref, n, r) if final else None, sig, bs, R, r,
                                  ps_num, ps_range, chroma, final, E, want_groups=True,
                                  options=(COSTS.index(cost), 0, 0))
        out.append(res)
        groups.append(g)
        count += c
    return np.stack(out), np.stack(groups), count


def _basic_from_stacks(stacks, noisy, r):
    N, P = noisy.shape[:2]
    basic = np.empty_like(noisy)
    for n in range(N):
        if r == 0:
            with np.errstate(invalid="ignore", divide="ignore"):
                basic[n] = stacks[n, :, 0, 0] / stacks[n, :, 0, 1]
        else:
            win = [stacks[min(max(n - r + i, 0), N - 1)] for i in range(2 * r + 1)]
            basic[n] = O.vaggregate(win, noisy[n], n, N, r, [True] * P)
    return np.where(np.isfinite(basic), basic, noisy).astype(np.float32)


@pytest.mark.parametrize("extractor_exp", [10, 0])
@pytest.mark.parametrize("name", list(CASES))
def test_reach_two_pass_stacks(name, extractor_exp):
    core = _core()
    case = CASES[name]
    H, W, N, P, chroma, bs, R, r, ps_num, ps_range, cost, reduced = case
    sigma = [12.0] * P
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=12.0, seed=200 + R + r, motion=True, yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    E = float(2 ** extractor_exp) if extractor_exp else 0.0

    def plan(final):
        return core.Plan(W, H, sigma, bs, R, r, ps_num, ps_range, chroma=chroma, final=final,
                         extractor_exp=extractor_exp, bm_error=cost)

    def check(got, want, what):
        if extractor_exp:
            np.testing.assert_array_equal(got, want, err_msg=what)
        else:
            np.testing.assert_allclose(got, want, rtol=2e-6, atol=5e-6, err_msg=what)

    pb = plan(False)
    if r > 0:
        # the accumulation margin is below the reach exactly when the fused
        # VAggregate route is refused
        if reduced:
            with pytest.raises(core.C.BM3DError, match="search reach too large for region aggregation"):
                pb.execute_vagg(src)
        else:
            pb.execute_vagg(src)
    sig_b = [O.scale_sigma(s, False) for s in sigma]
    want_b, want_g, want_n = _oracle(noisy, None, sig_b, case, False, E)
    got_g = core.unpack_groups(pb.match(src))
    np.testing.assert_array_equal(got_g, want_g)
    assert pb.count_candidates(src) == want_n
    got_b = pb.execute(src).cpu().numpy()
    check(got_b, want_b, "basic stacks")
    # final pass on the oracle's basic estimate (the same ref for both sides)
    basic = _basic_from_stacks(want_b, noisy, r)
    pf = plan(True)
    got_f = pf.execute(src, torch.from_numpy(basic).cuda()).cpu().numpy()
    sig_f = [O.scale_sigma(s, True) for s in sigma]
    want_f, _, _ = _oracle(noisy, basic, sig_f, case, True, E)
    check(got_f, want_f, "final stacks")


def test_reach_batch_refuses_plane_matcher():
    core = _core()
    H, W, N, P, chroma, bs, R, r, ps_num, ps_range, _, _ = CASES["R64-r0"]
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=12.0, seed=7)
    src = torch.from_numpy(noisy).cuda()
    job = dict(src=src, passes=2, sigma=[12.0], block_step=bs, bm_range=R, radius=r, ps_num=ps_num,
               ps_range=ps_range)
    with pytest.raises(core.C.BM3DError, match="search reach too large for a batch job"):
        core.Batch([job])


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_bm3dv2_radius4_falls_back_to_stacks(precision):
    """BM3Dv2 radius 4 with the reference's default ranges (bs 8, bm_range 9,
    ps_range 4): the filter's margin is reduced, so the plugin filter runs
    the stacks and VAggregate instead of the fused route."""
    _core()
    from bm3d_amd import vs
    H, W, N = 48, 64, 10
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=41, motion=True)
    src = vs.Clip.from_array(noisy, "GRAY")
    kw = dict(sigma=[15.0], block_step=8, bm_range=9, radius=4, ps_num=2, ps_range=4)
    prm = O.Params(sigma=(15.0,), block_step=8, bm_range=9, radius=4, ps_num=2, ps_range=4)
    if precision == "fp32":
        basic = vs.BM3Dv2(src, extractor_exp=10, **kw)
        final = vs.BM3Dv2(src, ref=basic, extractor_exp=10, **kw).numpy()
        ob, of = O.two_pass(noisy, dataclasses.replace(prm, extractor_exp=10))
        np.testing.assert_array_equal(basic.numpy(), ob)
        np.testing.assert_array_equal(final, of)
    else:
        basic = vs.BM3Dv2(src, precision="mixed", **kw)
        final = vs.BM3Dv2(src, ref=basic, precision="mixed", **kw).numpy()
        _, f64 = O.two_pass(noisy.astype(np.float64), prm, dtype=np.float64)
        d = synth.psnr(final, clean) - synth.psnr(f64, clean)
        assert abs(d) <= 0.05, d
    assert synth.psnr(final, clean) > synth.psnr(noisy, clean) + 3.0
