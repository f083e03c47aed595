"""Pins of the CPU oracle against what the mathematics fixes.

The reference ships no tests, fixtures or golden data, and its CPU core
cannot be built here (it needs the VapourSynth headers), so each oracle
function is checked against closed forms, invariants and independent
brute-force formulations chosen so that a plausible slip (wrong constant,
dropped scale, off-by-one window, wrong tie rule, wrong slot formula) fails:

* DCT: against the 4 x orthonormal DCT-II matrix built from cosines;
* distance: against a float64 sum of squares;
* top-8 / centre insert / predictive temporal search: against an order-free
  restatement that sorts candidates by (error, arrival index);
* shrinkage: sigma = 0 identity, DC-only group mean, Wiener limits;
* pipeline: constant image invariance, FP32 vs FP64 agreement;
* VAggregate: the slot formula's target-frame invariant and border cases.
"""
import numpy as np
import pytest

from bm3d_amd import synth
from oracle import oracle as O


def dct_matrix():
    """4 x orthonormal DCT-II: M[k][n] = 4 a_k cos(pi (2n+1) k / 16)."""
    k = np.arange(8)[:, None]
    n = np.arange(8)[None, :]
    a = np.where(k == 0, np.sqrt(1 / 8), np.sqrt(2 / 8))
    return 4.0 * a * np.cos(np.pi * (2 * n + 1) * k / 16)


# ------------------------------------------------------------------ DCT
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_dct8_matches_cosine_matrix(dtype, tol):
    rng = np.random.default_rng(1)
    x = rng.normal(size=(50, 8))
    M = dct_matrix()
    fwd = O.dct8(x, dtype=dtype)
    np.testing.assert_allclose(fwd, x @ M.T, rtol=0, atol=tol * 8)
    inv = O.dct8(x, inverse=True, dtype=dtype)
    np.testing.assert_allclose(inv, x @ M, rtol=0, atol=tol * 8)


def test_dct8_roundtrip_is_16x():
    x = np.random.default_rng(2).normal(size=(20, 8))
    y = O.dct8(O.dct8(x, dtype=np.float64), inverse=True, dtype=np.float64)
    np.testing.assert_allclose(y, 16 * x, atol=1e-12)


def test_transform3d_separable_and_4096():
    rng = np.random.default_rng(3)
    g = rng.normal(size=(8, 8, 8))
    M = dct_matrix()
    want = np.einsum("ai,bj,ck,ijk->abc", M, M, M, g)
    got = O.transform3d(g, dtype=np.float64)
    np.testing.assert_allclose(got, want, atol=1e-10)
    back = O.transform3d(got, inverse=True, dtype=np.float64)
    np.testing.assert_allclose(back, 4096 * g, atol=1e-8)
    c = O.transform3d(np.full((8, 8, 8), 0.25), dtype=np.float64)
    assert abs(c[0, 0, 0] - 0.25 * (8 * np.sqrt(2)) ** 3) < 1e-9
    c[0, 0, 0] = 0
    assert np.abs(c).max() < 1e-10


# ------------------------------------------------------------- distance
def test_ssd_matches_float64():
    rng = np.random.default_rng(4)
    img = rng.uniform(size=(20, 24)).astype(np.float32)
    ref = img[3:11, 5:13].copy()
    for (x, y) in [(0, 0), (5, 3), (16, 12), (9, 7)]:
        want = float(np.sum((ref.astype(np.float64) - img[y:y + 8, x:x + 8]) ** 2))
        assert abs(O.ssd(ref, img, x, y) - want) <= 1e-6 * max(want, 1e-3)
    assert O.ssd(ref, img, 5, 3) == 0.0


# ----------------------------------------------- selection (order-free)
def _brute_lists(slots, bs, R, radius, ps_num, ps_range):
    """Order-free restatement: every list is the 8 smallest keys
    (error, arrival index) of the candidates seen, with the placeholders
    (FLT_MAX, arrival -1) of cpu_source/source.cpp:802-812."""
    H, W = slots[0].shape
    BIG = float(np.finfo(np.float32).max)
    out = []
    ys = O.block_positions(H, bs)
    xs = O.block_positions(W, bs)
    for y in ys:
        for x in xs:
            ref = slots[radius][y:y + 8, x:x + 8]
            arrival = [0]

            def window(cx, cy, rng_, z, keys):
                for row in range(max(cy - rng_, 0), min(cy + rng_, H - 8) + 1):
                    for col in range(max(cx - rng_, 0), min(cx + rng_, W - 8) + 1):
                        e = O.ssd(ref, slots[z], col, row)
                        keys.append((e, arrival[0], col, row, z))
                        arrival[0] += 1

            def top8(keys):
                return sorted(keys, key=lambda k: (k[0], k[1]))[:8]

            L = [(BIG, -1, x, y, radius)] * 8
            keys = list(L)
            window(x, y, R, radius, keys)
            L = top8(keys)
            if radius:
                centre = [(k[2], k[3]) for k in L]
                for d in (-1, 1):
                    last = centre
                    for t in range(1, radius + 1):
                        z = radius + d * t
                        F = [(BIG, -1, 0, 0, z)] * 8
                        for i in range(ps_num):
                            F = top8(F + _collect(window, last[i], ps_range, z))
                        L = top8(L + [(k[0], arrival[0] + i, k[2], k[3], k[4])
                                      for i, k in enumerate(F[:ps_num])])
                        arrival[0] += ps_num
                        last = [(k[2], k[3]) for k in F]
            pos = [(k[2], k[3], k[4]) for k in L]
            if (x, y, radius) not in pos:
                pos = [(x, y, radius)] + pos[:7]
            out.append(pos)
    return np.array(out)


def _collect(window, c, rng_, z):
    keys = []
    window(c[0], c[1], rng_, z, keys)
    return keys


@pytest.mark.parametrize("quantise", [False, True])
def test_spatial_top8_order_free(quantise):
    rng = np.random.default_rng(5)
    img = rng.uniform(size=(22, 26)).astype(np.float32)
    if quantise:  # many exact ties
        img = (np.round(img * 2) / 2).astype(np.float32)
    g, _, n = O.match([img], bs=3, bm_range=3)
    want = _brute_lists([img], 3, 3, 0, 2, 4)
    np.testing.assert_array_equal(g, want)


@pytest.mark.parametrize("radius,ps_num,ps_range", [(1, 2, 2), (2, 3, 1), (1, 8, 1)])
def test_temporal_top8_order_free(radius, ps_num, ps_range):
    rng = np.random.default_rng(6 + radius)
    base = rng.uniform(size=(20, 22)).astype(np.float32)
    slots = [np.roll(base, (k, -k), axis=(0, 1)) + rng.normal(0, 0.02, base.shape).astype(np.float32)
             for k in range(2 * radius + 1)]
    slots = [np.ascontiguousarray(s, dtype=np.float32) for s in slots]
    g, _, _ = O.match(slots, bs=4, bm_range=2, radius=radius, ps_num=ps_num, ps_range=ps_range)
    want = _brute_lists(slots, 4, 2, radius, ps_num, ps_range)
    np.testing.assert_array_equal(g, want)


def test_temporal_duplicates_kept():
    # identical slots: overlapping ps windows see the same blocks twice and
    # the reference keeps both copies (README.md:80)
    img = np.random.default_rng(7).uniform(size=(16, 16)).astype(np.float32)
    g, e, _ = O.match([img, img, img], bs=8, bm_range=1, radius=1, ps_num=2, ps_range=1)
    want = _brute_lists([img, img, img], 8, 1, 1, 2, 1)
    np.testing.assert_array_equal(g, want)
    pos = [tuple(p) for p in g[0]]
    assert len(set(pos)) < len(pos)


def test_candidate_count_closed_form():
    H, W, bs, R = 40, 52, 4, 5
    img = np.random.default_rng(8).uniform(size=(H, W)).astype(np.float32)
    _, _, n = O.match([img], bs=bs, bm_range=R)
    want = sum((min(y + R, H - 8) - max(y - R, 0) + 1) * (min(x + R, W - 8) - max(x - R, 0) + 1)
               for y in O.block_positions(H, bs) for x in O.block_positions(W, bs))
    assert n == want


def test_block_positions():
    assert O.block_positions(8, 4) == [0]
    assert O.block_positions(17, 4) == [0, 4, 8, 9]
    assert O.block_positions(16, 8) == [0, 8]
    assert O.block_positions(20, 8) == [0, 8, 12]


# ------------------------------------------------------------ shrinkage
def test_hard_sigma_zero_is_identity():
    g = np.random.default_rng(9).uniform(size=(8, 8, 8))
    est, w = O.collab_hard(g, 0.0, dtype=np.float64)
    np.testing.assert_allclose(est, g, atol=1e-12)
    assert w == 1.0 / 512


def test_hard_huge_sigma_keeps_dc_only():
    g = np.random.default_rng(10).uniform(size=(8, 8, 8))
    est, w = O.collab_hard(g, 1e9, dtype=np.float64)
    np.testing.assert_allclose(est, np.full_like(g, g.mean()), atol=1e-12)
    assert w == 1.0


def test_hard_threshold_counts_kept():
    g = np.random.default_rng(11).normal(size=(8, 8, 8))
    c = O.transform3d(g, dtype=np.float64)
    thr = np.quantile(np.abs(c), 0.7)
    est, w = O.collab_hard(g, thr, dtype=np.float64)
    keep = np.abs(c) >= thr
    keep[0, 0, 0] = True
    assert w == 1.0 / keep.sum()
    M = dct_matrix()
    want = np.einsum("ai,bj,ck,abc->ijk", M, M, M, np.where(keep, c, 0.0)) / 4096
    np.testing.assert_allclose(est, want, atol=1e-10)


def test_wiener_limits():
    rng = np.random.default_rng(12)
    g = rng.uniform(size=(8, 8, 8))
    # basic estimate with huge coefficients everywhere -> coef ~ 1 -> identity
    est, w = O.collab_wiener(g, g * 1e6, 1e-3, dtype=np.float64)
    np.testing.assert_allclose(est, g, atol=1e-9)
    assert abs(w - 1.0 / 512) < 1e-12
    # constant basic estimate -> only DC coefficient nonzero -> group mean
    est, w = O.collab_wiener(g, np.full_like(g, 0.5), 1.0, dtype=np.float64)
    np.testing.assert_allclose(est, np.full_like(g, g.mean()), atol=1e-12)
    assert w == 1.0


def test_wiener_coefficient_formula():
    rng = np.random.default_rng(13)
    g, b = rng.uniform(size=(8, 8, 8)), rng.uniform(size=(8, 8, 8))
    s = 3.0
    est, w = O.collab_wiener(g, b, s, dtype=np.float64)
    cb, cg = O.transform3d(b, dtype=np.float64), O.transform3d(g, dtype=np.float64)
    coef = cb ** 2 / (cb ** 2 + s * s)
    coef[0, 0, 0] = 1.0
    M = dct_matrix()
    want = np.einsum("ai,bj,ck,abc->ijk", M, M, M, cg * coef) / 4096
    np.testing.assert_allclose(est, want, atol=1e-10)
    assert abs(w - 1.0 / np.sum(coef ** 2)) < 1e-14


# ------------------------------------------------------------- pipeline
@pytest.mark.parametrize("radius", [0, 1])
def test_constant_image_is_fixed_point(radius):
    c = np.full((3, 1, 24, 40), 0.3, np.float32)
    p = O.Params(sigma=(10.0,), block_step=4, bm_range=4, radius=radius, ps_num=2, ps_range=2)
    basic, final = O.two_pass(c, p)
    assert np.abs(basic - c).max() <= 2e-7
    assert np.abs(final - c).max() <= 2e-7


def test_fp32_close_to_fp64():
    _, noisy = synth.make_clip(40, 48, 3, 1, noise_sd=15, seed=2, motion=True)
    p = O.Params(sigma=(15.0,), block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2)
    b32, f32 = O.two_pass(noisy, p)
    b64, f64 = O.two_pass(noisy.astype(np.float64), p, dtype=np.float64)
    assert np.abs(b32 - b64).max() < 1e-5
    assert np.abs(f32 - f64).max() < 1e-4


def test_denoising_improves_psnr():
    clean, noisy = synth.make_clip(64, 96, 3, 1, noise_sd=20, seed=3, motion=True)
    p = O.Params(sigma=(20.0,), block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3)
    basic, final = O.two_pass(noisy, p)
    assert synth.psnr(basic, clean) > synth.psnr(noisy, clean) + 3
    assert synth.psnr(final, clean) > synth.psnr(noisy, clean) + 3


def test_chroma_skips_zero_sigma_plane():
    clean, noisy = synth.make_clip(32, 40, 1, 3, noise_sd=10, seed=4, yuv=True)
    p = O.Params(sigma=(10.0, 0.0, 10.0), block_step=4, bm_range=4, chroma=True)
    out = O.bm3dv2(noisy, None, p)
    np.testing.assert_array_equal(out[:, 1], noisy[:, 1])
    assert np.abs(out[:, 0] - noisy[:, 0]).max() > 0


# ------------------------------------------------------------ VAggregate
@pytest.mark.parametrize("radius,N", [(1, 5), (2, 6), (1, 1), (2, 3)])
def test_vaggregate_target_invariant(radius, N):
    """Slot z of frame m's stack holds contributions to frame m - r + z;
    with num = target id and den = 1 every interior output is its own id."""
    H, W = 2, 3
    tw = 2 * radius + 1
    stacks = []
    for m in range(N):
        s = np.zeros((1, tw, 2, H, W), np.float32)
        for z in range(tw):
            s[0, z, 0] = 100 * m + z
            s[0, z, 1] = 1
        stacks.append(s)
    for n in range(N):
        win = [stacks[min(max(n - radius + i, 0), N - 1)] for i in range(tw)]
        out = O.vaggregate(win, np.zeros((1, H, W), np.float32), n, N, radius, [True])
        # hand evaluation of source/source.cpp:834-836
        tot = 0.0
        for i in range(tw):
            f = min(max(n - radius + i, 0), N - 1)
            s = min(max(2 * radius - i, n - N + 1 + radius), n + radius)
            tot += 100 * f + s
        assert np.allclose(out, tot / tw)
        if radius <= n < N - radius:
            # interior: every term is 100 (n - r + i) + (2r - i) -> mean 100 n + r - 99 r ... -> target
            targets = [(n - radius + i) + (2 * radius - i) - radius for i in range(tw)]
            assert all(t == n for t in targets)


def test_vaggregate_border_example():
    """n = 0, r = 1, N = 5: stack(0) slot 1 twice plus stack(1) slot 0."""
    stacks = []
    for m in range(5):
        s = np.zeros((1, 3, 2, 1, 1), np.float32)
        for z in range(3):
            s[0, z, 0] = 10 * m + z
            s[0, z, 1] = 1
        stacks.append(s)
    win = [stacks[0], stacks[0], stacks[1]]
    out = O.vaggregate(win, np.zeros((1, 1, 1), np.float32), 0, 5, 1, [True])
    assert out[0, 0, 0] == np.float32((1 + 1 + 10) / 3)
