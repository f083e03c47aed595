"""Headroom of the fixed-point accumulation.

The filters add weighted estimates into int32 LDS regions in units of
2^(e-24) (extractor_exp = e) or 2^-22 scaled by the input's range. A pixel's
sum can exceed 2^31 units: on flat, noise-free content every group keeps only
its DC coefficient (weight 1) and, with block_step 1, every pixel receives
about 512 members of weight 1 per slot (ties keep the first positions of each
search window, and every position is a member of about 8 groups), twice that
at the clip's ends where VAggregate reads a slot twice. The kernels track
every 32-bit wrap of a region word exactly and add the wraps back when the
regions are read, so these sums match the oracle's float sums.

Cases: constant 1.0 and 100.0 clips at block_step 1 and 2 (sums of 512 to
about 150000 x the unit's headroom), radius 0 and 1, extractor off and
small; and extractor_exp = 1 on data scaled to [0, 255], where a single
weighted estimate exceeds 2^30 units and is split into 32-bit words.
"""
import numpy as np
import pytest

from bm3d_amd import synth
from oracle import oracle as O

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import core
    return core


def _window(clip, n, r):
    N = clip.shape[0]
    return np.stack([clip[min(max(n - r + z, 0), N - 1)] for z in range(2 * r + 1)], axis=1)


def _oracle(noisy, ref, prm, final):
    N, r = noisy.shape[0], prm["radius"]
    sig = [O.scale_sigma(prm["sigma"], final)]
    E = float(2.0 ** prm["extractor_exp"]) if prm["extractor_exp"] else 0.0
    return np.stack([O.bm3d_window(_window(noisy, n, r), _window(ref, n, r) if final else None, sig,
                                   prm["block_step"], prm["bm_range"], r, prm["ps_num"], prm["ps_range"],
                                   False, final, E) for n in range(N)])


def _oracle_out(stacks, noisy, r):
    N = stacks.shape[0]
    if r == 0:
        return stacks[:, :, 0, 0] / stacks[:, :, 0, 1]
    return np.stack([O.vaggregate([stacks[min(max(n - r + i, 0), N - 1)] for i in range(2 * r + 1)],
                                  noisy[n], n, N, r, [True]) for n in range(N)])


def _check(core, clip, prm, precision, exact, rtol=2e-6, out_atol=None):
    H, W = clip.shape[-2:]
    src = torch.from_numpy(clip).cuda()
    r = prm["radius"]
    ref = None
    for final in (False, True):
        plan = core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], r, prm["ps_num"],
                         prm["ps_range"], final=final, extractor_exp=prm["extractor_exp"], precision=precision)
        rt = torch.from_numpy(ref).cuda() if final else None
        got = plan.execute(src, rt).cpu().numpy()
        fused = plan.execute_vagg(src, rt).cpu().numpy()
        want = _oracle(clip, ref, prm, final)
        want_out = _oracle_out(want, clip, r).astype(np.float32)
        tag = f"{precision} {'final' if final else 'basic'}"
        if exact:
            np.testing.assert_array_equal(got, want, err_msg=tag)
            np.testing.assert_array_equal(fused, want_out, err_msg=tag)
        elif out_atol is not None:
            # outputs only: FP16 transforms can flip hard-threshold decisions,
            # which changes the weights in the stacks
            assert np.isfinite(fused).all(), tag
            assert np.abs(fused - want_out).max() <= out_atol, tag
        else:
            scale = float(np.abs(clip).max())
            np.testing.assert_allclose(got, want, rtol=rtol, atol=5e-6 * scale, err_msg=tag)
            np.testing.assert_allclose(fused, want_out, rtol=rtol, atol=5e-6 * scale, err_msg=tag)
        ref = want_out


FLAT = [
    # (value, block_step, radius, extractor_exp)
    (1.0, 1, 0, 0),
    (1.0, 1, 1, 0),
    (1.0, 2, 1, 1),
    (1.0, 1, 1, 3),
    (100.0, 1, 1, 0),
]


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("value,bs,r,e", FLAT, ids=[f"v{v:g}-bs{b}-r{r}-e{e}" for v, b, r, e in FLAT])
def test_flat_clip_sums_beyond_int32(value, bs, r, e, precision):
    """Constant clips are fixed points of both passes; their per-pixel sums
    pass 2^31 units, and stacks and outputs still equal the oracle's bit for
    bit (every term is a whole number of units, so every sum is exact)."""
    core = _gpu()
    clip = np.full((4, 1, 40, 48), value, np.float32)
    prm = dict(sigma=10.0, block_step=bs, bm_range=4, radius=r, ps_num=2, ps_range=2, extractor_exp=e)
    _check(core, clip, prm, precision, exact=e > 0)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_single_estimates_beyond_2_30_units(precision):
    """extractor_exp = 1 on data in [0, 255]: one weighted estimate is up to
    255 / 2^-23 > 2^30 units, so every value takes the split 64-bit path.
    The oracle's float sums are then no longer exact, hence the tolerance
    (FP32); the mixed filter is held to its FP16 bound on the outputs (a
    flipped coefficient c moves a pixel by at most 8 |c| / 4096)."""
    core = _gpu()
    _, noisy = synth.make_clip(32, 40, 3, 1, noise_sd=10, seed=21, motion=True)
    clip = (noisy * np.float32(255.0)).astype(np.float32)
    prm = dict(sigma=10.0 * 255.0, block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=1)
    if precision == "fp32":
        _check(core, clip, prm, precision, exact=False, rtol=1e-5)
    else:
        _check(core, clip, prm, precision, exact=False, out_atol=0.02 * 255.0)
