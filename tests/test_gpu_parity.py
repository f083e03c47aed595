"""GPU parity of the HIP path against the CPU oracle (FP32 mirror).

Bars:
* group lists (block matching, temporal search, centre insert): identical;
* stacks with extractor_exp set: bit-identical (the pre-rounded atomic sums
  are exact, so their order does not matter);
* stacks without extractor_exp: weighted sums accumulate in 2^-22 fixed
  point and weights in float atomics of unspecified order, so stacks agree
  to rtol 2e-6 + atol 5e-6 (32 contributions x 2^-23 rounding each).
"""
import numpy as np
import pytest

import mixed_ties
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


def _oracle_stacks(noisy, ref, prm, final, chroma, dtype=np.float32):
    """Oracle stacks for every frame: [N, planes, 2r+1, 2, H, W]."""
    N = noisy.shape[0]
    r = prm["radius"]
    sig = [O.scale_sigma(s, final) for s in prm["sigma"]]
    E = float(np.ldexp(np.float32(1), prm["extractor_exp"])) if prm["extractor_exp"] else 0.0
    out = []
    groups = []
    for n in range(N):
        res, g, _ = O.bm3d_window(_window(noisy, n, r), _window(ref, n, r) if final else None, sig,
                                  prm["block_step"], prm["bm_range"], r, prm["ps_num"],
                                  prm["ps_range"], chroma, final, E, dtype=dtype, want_groups=True)
        out.append(res)
        groups.append(g)
    return np.stack(out), np.stack(groups)


def _plan(core, W, H, prm, final, chroma):
    return core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], prm["radius"],
                     prm["ps_num"], prm["ps_range"], chroma=chroma, final=final,
                     extractor_exp=prm["extractor_exp"])


CASES = [
    # (H, W, frames, planes, chroma, params)
    (64, 80, 1, 1, False, dict(sigma=[10.0], block_step=4, bm_range=5, radius=0, ps_num=2,
                               ps_range=4, extractor_exp=0)),
    (72, 88, 1, 1, False, dict(sigma=[25.0], block_step=8, bm_range=9, radius=0, ps_num=2,
                               ps_range=4, extractor_exp=0)),
    (45, 61, 1, 1, False, dict(sigma=[15.0], block_step=3, bm_range=4, radius=0, ps_num=2,
                               ps_range=4, extractor_exp=0)),
    (40, 56, 4, 1, False, dict(sigma=[15.0], block_step=4, bm_range=4, radius=1, ps_num=2,
                               ps_range=3, extractor_exp=0)),
    (33, 47, 5, 1, False, dict(sigma=[20.0], block_step=5, bm_range=3, radius=2, ps_num=3,
                               ps_range=2, extractor_exp=0)),
    (40, 48, 3, 3, True, dict(sigma=[10.0, 8.0, 12.0], block_step=4, bm_range=4, radius=1,
                              ps_num=2, ps_range=2, extractor_exp=0)),
    (40, 48, 1, 3, True, dict(sigma=[10.0, 0.0, 12.0], block_step=6, bm_range=5, radius=0,
                              ps_num=2, ps_range=2, extractor_exp=0)),
]


def _ids():
    return [f"{h}x{w}x{n}f-p{p}{'c' if c else ''}-bs{q['block_step']}-R{q['bm_range']}-r{q['radius']}"
            for h, w, n, p, c, q in CASES]


@pytest.mark.parametrize("case", CASES, ids=_ids())
def test_match_groups_identical(case):
    core = _gpu()
    H, W, N, P, chroma, prm = case
    clean, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=3, motion=True,
                                   yuv=chroma)
    plan = _plan(core, W, H, prm, False, chroma)
    g = plan.match(torch.from_numpy(noisy).cuda())
    torch.cuda.synchronize()
    got = core.unpack_groups(g)
    _, want = _oracle_stacks(noisy, None, prm, False, chroma)
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("extractor_exp", [0, 12])
@pytest.mark.parametrize("case", CASES, ids=_ids())
def test_two_pass_stacks(case, extractor_exp):
    core = _gpu()
    H, W, N, P, chroma, prm = case
    prm = dict(prm, extractor_exp=extractor_exp)
    clean, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=5, motion=True,
                                   yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    # basic pass
    plan_b = _plan(core, W, H, prm, False, chroma)
    got_b = plan_b.execute(src).cpu().numpy()
    want_b, _ = _oracle_stacks(noisy, None, prm, False, chroma)
    # stacks of skipped chroma planes stay zero in both
    if extractor_exp:
        np.testing.assert_array_equal(got_b, want_b)
    else:
        np.testing.assert_allclose(got_b, want_b, rtol=2e-6, atol=5e-6)
    # final pass on the oracle's basic estimate (same ref for both sides)
    basic = np.empty_like(noisy)
    for n in range(N):
        if prm["radius"] == 0:
            num, den = want_b[n, :, 0, 0], want_b[n, :, 0, 1]
            with np.errstate(invalid="ignore", divide="ignore"):
                basic[n] = num / den
        else:
            win = [want_b[min(max(n - prm["radius"] + i, 0), N - 1)]
                   for i in range(2 * prm["radius"] + 1)]
            basic[n] = O.vaggregate(win, noisy[n], n, N, prm["radius"], [True] * P)
    basic = np.where(np.isfinite(basic), basic, noisy).astype(np.float32)
    plan_f = _plan(core, W, H, prm, True, chroma)
    got_f = plan_f.execute(src, torch.from_numpy(basic).cuda()).cpu().numpy()
    want_f, _ = _oracle_stacks(noisy, basic, prm, True, chroma)
    if extractor_exp:
        np.testing.assert_array_equal(got_f, want_f)
    else:
        np.testing.assert_allclose(got_f, want_f, rtol=2e-6, atol=5e-6)


@pytest.mark.parametrize("kind", ["x255", "offset1000"])
def test_stacks_wide_range_match_oracle(kind):
    """The fixed-point unit of the region accumulation follows the input's
    range (largest finite |sample| of the call), so inputs far outside
    [0, 1] -- which a fixed 2^-22 unit would overflow -- match the oracle's
    float sums to the same relative tolerance."""
    core = _gpu()
    H, W, N = 40, 56, 3
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=17, motion=True)
    scale, off = (255.0, 0.0) if kind == "x255" else (1.0, 1000.0)
    x = (noisy * np.float32(scale) + np.float32(off)).astype(np.float32)
    prm = dict(sigma=[15.0 * scale], block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=3,
               extractor_exp=0)
    plan = _plan(core, W, H, prm, False, False)
    got = plan.execute(torch.from_numpy(x).cuda()).cpu().numpy()
    want, _ = _oracle_stacks(x, None, prm, False, False)
    np.testing.assert_allclose(got, want, rtol=2e-6, atol=5e-6 * max(scale, off))


def test_vaggregate_and_aggregate_match_oracle():
    core = _gpu()
    H, W, N = 24, 40, 5
    rng = np.random.default_rng(0)
    r = 2
    stacks = rng.uniform(0.1, 1.0, size=(N, 1, 2 * r + 1, 2, H, W)).astype(np.float32)
    got = core.vaggregate(torch.from_numpy(stacks).cuda(), r, N).cpu().numpy()
    for n in range(N):
        win = [stacks[min(max(n - r + i, 0), N - 1)] for i in range(2 * r + 1)]
        want = O.vaggregate(win, np.zeros((1, H, W), np.float32), n, N, r, [True])
        np.testing.assert_array_equal(got[n], want)
    st0 = stacks[:, :, :1]
    got0 = core.aggregate(torch.from_numpy(np.ascontiguousarray(st0)).cuda()).cpu().numpy()
    np.testing.assert_array_equal(got0[:, 0], st0[:, 0, 0, 0] / st0[:, 0, 0, 1])


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_deterministic_with_extractor(precision):
    core = _gpu()
    H, W, N = 64, 96, 3
    prm = dict(sigma=[15.0], block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3,
               extractor_exp=8)
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=9, motion=True)
    src = torch.from_numpy(noisy).cuda()
    for final in (False, True):
        plan = core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], prm["radius"],
                         prm["ps_num"], prm["ps_range"], final=final, extractor_exp=8,
                         precision=precision)
        ref = src if final else None
        a = plan.execute(src, ref).cpu().numpy()
        b = plan.execute(src, ref).cpu().numpy()
        assert synth.sha256(a) == synth.sha256(b)
        c = plan.execute_vagg(src, ref).cpu().numpy()
        d = plan.execute_vagg(src, ref).cpu().numpy()
        assert synth.sha256(c) == synth.sha256(d)


# ---------------------------------------------------- mixed-precision filter
# The mixed filter evaluates the 3-D DCT pair as FP16 matrix products with
# FP32 accumulation: block pixels and intermediate coefficients are rounded
# to FP16 (relative 2^-11); shrinkage, weights and aggregation stay FP32.
# On the same group lists its per-slot estimates num/den agree with the FP32
# filter's to that rounding, apart from coefficients that sit on the hard
# threshold: a flipped coefficient c moves a pixel by at most 8 |c| / 4096
# (0.015 at sigma 15), which bounds the max-abs tolerance below.

def _slot_estimates(stack):
    num, den = stack[:, :, :, 0], stack[:, :, :, 1]
    with np.errstate(invalid="ignore", divide="ignore"):
        return num / den, den > 0


@pytest.mark.parametrize("extractor_exp", [0, 10])
@pytest.mark.parametrize("case", CASES, ids=_ids())
def test_mixed_filter_close_to_fp32(case, extractor_exp):
    core = _gpu()
    H, W, N, P, chroma, prm = case
    prm = dict(prm, extractor_exp=extractor_exp)
    r = prm["radius"]
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=5, motion=True, yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    basic = None
    # empirical bounds with headroom; threshold flips grow with sigma
    max_abs = 0.03 * max(prm["sigma"]) / 15.0
    max_rms = 1e-3 * max(prm["sigma"]) / 15.0
    for final in (False, True):
        p32 = _plan(core, W, H, prm, final, chroma)
        pmx = core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], r, prm["ps_num"],
                        prm["ps_range"], chroma=chroma, final=final, extractor_exp=extractor_exp,
                        precision="mixed")
        assert pmx.uses_mixed_filter()
        pmx.select_kernels("fp32", "mixed")  # same group lists as the FP32 plan
        assert pmx.uses_mixed_filter() and not pmx.uses_mixed_kernel()
        s32 = p32.execute(src, basic).cpu().numpy()
        smx = pmx.execute(src, basic).cpu().numpy()
        e32, m32 = _slot_estimates(s32)
        emx, mmx = _slot_estimates(smx)
        np.testing.assert_array_equal(mmx, m32)
        d = np.abs(emx - e32)[m32]
        rms = float(np.sqrt(np.mean(d.astype(np.float64) ** 2)))
        print(f"final={final}: rms {rms:.2e} max {d.max():.2e}")
        assert rms < max_rms and d.max() < max_abs
        if not final:
            est = core.vaggregate(torch.from_numpy(s32).cuda(), r, N) if r > 0 else \
                core.aggregate(torch.from_numpy(s32).cuda())
            est = est.cpu().numpy()
            basic = torch.from_numpy(np.where(np.isfinite(est), est, noisy).astype(np.float32)).cuda()


@pytest.mark.parametrize("kind", ["x255", "offset1000"])
def test_mixed_filter_wide_range(kind):
    """Inputs far outside [0, 1] (a 0-255 scale with sigma scaled alike, or a
    large offset): the mixed filter works on each group's residual about an
    offset, scaled by a power of two, so FP16 neither overflows nor loses
    the residual's precision. Both filters are compared with the FP64 oracle
    on the same input: the mixed filter's error stays at its FP16 level
    relative to the scale (and, with the offset, below the FP32 filter's,
    whose small coefficients lose digits to the offset)."""
    core = _gpu()
    H, W, N = 48, 64, 3
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=13, motion=True)
    scale, off = (255.0, 0.0) if kind == "x255" else (1.0, 1000.0)
    x = (noisy * np.float32(scale) + np.float32(off)).astype(np.float32)
    src = torch.from_numpy(x).cuda()
    prm = dict(block_step=4, bm_range=5, radius=1, ps_num=2, ps_range=3)
    oprm = dict(prm, sigma=[15.0 * scale], extractor_exp=0)
    basic = None
    for final in (False, True):
        p32 = core.Plan(W, H, 15.0 * scale, final=final, **prm)
        pmx = core.Plan(W, H, 15.0 * scale, final=final, precision="mixed", **prm)
        pmx.select_kernels("fp32", "mixed")
        s32 = p32.execute(src, basic).cpu().numpy()
        smx = pmx.execute(src, basic).cpu().numpy()
        s64, _ = _oracle_stacks(x, None if basic is None else basic.cpu().numpy(), oprm, final, False,
                                dtype=np.float64)
        e32, m32 = _slot_estimates(s32)
        emx, mmx = _slot_estimates(smx)
        e64, m64 = _slot_estimates(s64)
        np.testing.assert_array_equal(mmx, m32)
        assert np.isfinite(emx[mmx]).all()
        m = m32 & m64
        err32 = float(np.sqrt(np.mean((e32[m] - e64[m]) ** 2))) / scale
        errmx = float(np.sqrt(np.mean((emx[m] - e64[m]) ** 2))) / scale
        print(f"{kind} final={final}: rms vs fp64: fp32 {err32:.2e} mixed {errmx:.2e}")
        assert errmx < max(1e-3, 1.5 * err32)
        if not final:
            basic = core.vaggregate(torch.from_numpy(s32).cuda(), 1, N)


def test_kernel_selection():
    core = _gpu()
    p32 = core.Plan(64, 48, 10.0)
    assert not p32.uses_mixed_kernel() and not p32.uses_mixed_filter()
    with pytest.raises(core.C.BM3DError, match="precision MIXED"):
        p32.select_kernels("fp32", "mixed")
    p32.select_kernels("fp32", "fp32")
    pmx = core.Plan(64, 48, 10.0, precision="mixed")
    assert pmx.uses_mixed_kernel() and pmx.uses_mixed_filter()
    pmx.select_kernels("mixed", "fp32")
    assert pmx.uses_mixed_kernel() and not pmx.uses_mixed_filter()
    with pytest.raises(ValueError):
        pmx.select_kernels("fp16", "fp32")
    wide = core.Plan(64, 48, 10.0, bm_range=40, precision="mixed")
    assert not wide.uses_mixed_kernel() and wide.uses_mixed_filter()
    with pytest.raises(core.C.BM3DError, match="too large"):
        wide.select_kernels("mixed", "mixed")


# ------------------------------------------------ BM3D + VAggregate fused
@pytest.mark.parametrize("extractor_exp", [10, 0])
@pytest.mark.parametrize("case", [
    (48, 64, 5, 1, False, dict(sigma=[12.0], block_step=4, bm_range=5, radius=1, ps_num=2, ps_range=3)),
    (40, 56, 6, 1, False, dict(sigma=[15.0], block_step=3, bm_range=4, radius=2, ps_num=2, ps_range=2)),
    (33, 47, 3, 1, False, dict(sigma=[20.0], block_step=5, bm_range=3, radius=2, ps_num=3, ps_range=2)),
    (40, 48, 4, 3, True, dict(sigma=[10.0, 0.0, 12.0], block_step=4, bm_range=4, radius=1, ps_num=2,
                              ps_range=2)),
], ids=["r1", "r2", "r2-3frames", "chroma-r1"])
def test_execute_vagg_equals_stacks_then_vaggregate(case, extractor_exp):
    """bm3d_execute_vagg gives the values of bm3d_execute + bm3d_vaggregate,
    for the whole clip and for frame subranges (evaluations of neighbouring
    frames are recomputed at the range edges)."""
    core = _gpu()
    H, W, N, P, chroma, prm = case
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=5, motion=True, yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    plan = core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], prm["radius"],
                     prm["ps_num"], prm["ps_range"], chroma=chroma, extractor_exp=extractor_exp)
    want = core.vaggregate(plan.execute(src), prm["radius"], N).cpu().numpy()
    keep = [p for p in range(P) if not (chroma and prm["sigma"][p] == 0.0)]
    for first, count in [(0, N), (1, N - 1), (N - 1, 1), (N // 2, 1)]:
        got = plan.execute_vagg(src, first=first, count=count).cpu().numpy()
        w = want[first:first + count][:, keep]
        g = got[:, keep]
        if extractor_exp:
            np.testing.assert_array_equal(g, w)
        else:
            np.testing.assert_allclose(g, w, rtol=2e-6, atol=5e-6)


# ---------------------------------------------------------- mixed precision
# The mixed path pre-scores candidates in FP16 and rescores the best K
# exactly (K = 10 spatial; 4, 8 or 16 per temporal frame for ps_num up to 2,
# 4 or 8); its lists equal the FP32 path's except where an exact top-k
# candidate ranks below K approximately (in practice at near ties). Slots too wide for FP16 (or not finite) are
# searched exactly, so those inputs must match bit for bit.

def _mixed_input(kind, noisy):
    if kind == "noisy":
        return noisy
    if kind == "quantised":
        return (np.round(noisy * 8.0) / 8.0).astype(np.float32)
    if kind == "flat":
        return np.full_like(noisy, 0.25)
    if kind == "wide":
        return (noisy * 40.0).astype(np.float32)
    if kind == "offset":
        return (noisy + 1000.0).astype(np.float32)
    if kind == "nan":
        out = noisy.copy()
        out[..., out.shape[-2] // 2, out.shape[-1] // 3] = np.nan
        return out
    raise ValueError(kind)


# input kinds; True: the lists must be identical (no ties, or the FP32 fallback)
MIXED_KINDS = {"noisy": False, "quantised": False, "flat": True, "wide": True, "offset": False, "nan": True}


@pytest.mark.parametrize("kind", list(MIXED_KINDS))
@pytest.mark.parametrize("case", CASES, ids=_ids())
def test_mixed_groups_agree_fp32(case, kind):
    """Mixed lists equal the FP32 lists except among exactly tied candidates
    (quantised and offset inputs have many exact ties)."""
    core = _gpu()
    H, W, N, P, chroma, prm = case
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=3, motion=True, yuv=chroma)
    x = _mixed_input(kind, noisy)
    src = torch.from_numpy(x).cuda()
    p32 = _plan(core, W, H, prm, False, chroma)
    pmx = core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], prm["radius"],
                    prm["ps_num"], prm["ps_range"], chroma=chroma, precision="mixed")
    assert pmx.uses_mixed_kernel()
    g32 = core.unpack_groups(p32.match(src))
    gmx = core.unpack_groups(pmx.match(src))
    if MIXED_KINDS[kind]:
        np.testing.assert_array_equal(gmx, g32)
    else:
        nd, nl = mixed_ties.tie_only(gmx, g32, x, prm["radius"], prm["block_step"])
        print(f"{kind}: {nd} of {nl} lists differ, all at exact ties")
    assert pmx.count_candidates(src) == p32.count_candidates(src)


@pytest.mark.parametrize("prm", [
    dict(block_step=4, bm_range=9, radius=1, ps_num=2, ps_range=4),
    dict(block_step=2, bm_range=12, radius=2, ps_num=3, ps_range=5),
    dict(block_step=7, bm_range=6, radius=1, ps_num=8, ps_range=1),
    dict(block_step=1, bm_range=2, radius=3, ps_num=1, ps_range=2),
    dict(block_step=3, bm_range=9, radius=1, ps_num=2, ps_range=4),
    dict(block_step=6, bm_range=9, radius=1, ps_num=4, ps_range=3),
], ids=["bench", "bs2-R12-r2", "bs7-ps8", "bs1-r3", "bs3", "bs6-ps4"])
def test_mixed_groups_agree_fp32_larger(prm):
    """Interior windows take the matcher's unchecked scan, border windows the
    checked one; block steps 3, 6 and 7 run its second kernel variant (block
    step not dividing 8). Bar: at most 0.01% of the lists differ from the
    FP32 path's, and only among near-tied candidates (exact distances within
    2^-7 of each other): the FP16 pre-score can rank such candidates either
    way, and the kept keys are a margin, not a certificate. The bench
    workload's lists are held to exact ties only (test_gpu_workloads.py)."""
    core = _gpu()
    H, W, N = 200, 264, 5
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=11, motion=True)
    src = torch.from_numpy(noisy).cuda()
    p32 = core.Plan(W, H, 15.0, **prm)
    pmx = core.Plan(W, H, 15.0, precision="mixed", **prm)
    assert pmx.uses_mixed_kernel()
    basic = core.vaggregate(p32.execute(src), prm["radius"], N)
    for final in (False, True):
        if final:
            p32 = core.Plan(W, H, 15.0, final=True, **prm)
            pmx = core.Plan(W, H, 15.0, final=True, precision="mixed", **prm)
            g32 = core.unpack_groups(p32.match(src, basic))
            gmx = core.unpack_groups(pmx.match(src, basic))
            matched = basic.cpu().numpy()
        else:
            g32 = core.unpack_groups(p32.match(src))
            gmx = core.unpack_groups(pmx.match(src))
            matched = noisy
        nd, nl = mixed_ties.tie_only(gmx, g32, matched, prm["radius"], prm["block_step"], rel=2.0 ** -7)
        print(f"final={final}: {nd} of {nl} lists differ, all at near-ties")
        assert nd <= 1e-4 * nl


def test_mixed_temporal_keys_option():
    """desc.mixed_temporal_keys: keeping more candidates per temporal frame
    for exact rescoring keeps a superset, so every list that matched the FP32
    path's still does; 0 means 8 at ps_num 2, whose lists differ from the
    FP32 path's only at exact ties. Values other than 0, 4, 8, 16, or below
    2 ps_num, are rejected."""
    core = _gpu()
    H, W, N = 200, 264, 5
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=12, motion=True)
    src = torch.from_numpy(noisy).cuda()
    prm = dict(block_step=4, bm_range=9, radius=1, ps_num=2, ps_range=4)
    g32 = core.Plan(W, H, 15.0, **prm).match(src).cpu().numpy()
    lists, agree = {}, {}
    for kt in (0, 4, 8, 16):
        pmx = core.Plan(W, H, 15.0, precision="mixed", mixed_temporal_keys=kt, **prm)
        assert pmx.uses_mixed_kernel()
        lists[kt] = pmx.match(src).cpu().numpy()
        agree[kt] = np.all(lists[kt] == g32, axis=-1)
    print({k: float(v.mean()) for k, v in agree.items()})
    np.testing.assert_array_equal(lists[0], lists[8])
    assert not np.any(agree[4] & ~agree[8]) and not np.any(agree[8] & ~agree[16])
    mixed_ties.tie_only(core.unpack_groups(torch.from_numpy(lists[8])), core.unpack_groups(torch.from_numpy(g32)),
                        noisy, prm["radius"], prm["block_step"])
    for bad, pn in [(6, 2), (32, 2), (4, 3), (-4, 1)]:
        with pytest.raises(Exception, match="mixed_temporal_keys"):
            core.Plan(W, H, 15.0, precision="mixed", mixed_temporal_keys=bad, **dict(prm, ps_num=pn))


def test_mixed_two_pass_within_005db_of_fp64():
    """The mixed path's quality bar: two-pass V-BM3D through it is within
    0.05 dB PSNR of the FP64 oracle running the same pipeline."""
    core = _gpu()
    H, W, N = 96, 128, 4
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=21, motion=True)
    prm = dict(sigma=15.0, block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3)
    src = torch.from_numpy(noisy).cuda()
    out = {}
    for prec in ("fp32", "mixed"):
        pb = core.Plan(W, H, precision=prec, **prm)
        pf = core.Plan(W, H, final=True, precision=prec, **prm)
        basic = core.vaggregate(pb.execute(src), 1, N)
        out[prec] = core.vaggregate(pf.execute(src, basic), 1, N).cpu().numpy()
    p = O.Params(sigma=(15.0,), block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3)
    _, f64 = O.two_pass(noisy.astype(np.float64), p, dtype=np.float64)
    ref_psnr = synth.psnr(f64, clean)
    for prec, o in out.items():
        assert abs(synth.psnr(o, clean) - ref_psnr) <= 0.05, prec
        assert np.abs(o - f64).max() < 0.05, prec
