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


def _ssd32(r, c):
    return O.ssd(r.astype(np.float32), c.astype(np.float32), 0, 0)


def test_ssd_summation_structure():
    """The FP32 distance follows cpu_source/source.cpp:98-103,123-136: per
    column an FMA chain over the even rows and one over the odd rows, their
    sum, then ((s0+s1)+(s2+s3))+((s4+s5)+(s6+s7)). That order is bitwise
    invariant under exactly the swaps that permute its own additions:
    exchanging the even and odd rows (the two chains), exchanging columns
    2c and 2c+1, the column pairs within a half, and the two halves. Orders a
    slip would plausibly produce (one chain over all rows, row-major sums, a
    sequential column sum) fail these on random data; swaps outside that
    group (rows 0 and 2) must change some result, so the chain really is
    sequential."""
    rng = np.random.default_rng(41)
    rows_eo = [1, 0, 3, 2, 5, 4, 7, 6]
    cols = {"pairs": [1, 0, 3, 2, 5, 4, 7, 6], "quads": [2, 3, 0, 1, 6, 7, 4, 5],
            "halves": [4, 5, 6, 7, 0, 1, 2, 3]}
    changed_rows02 = changed_cols12 = 0
    for _ in range(400):
        scale = 2.0 ** rng.integers(-6, 3)
        r = rng.uniform(-1, 1, (8, 8)) * scale
        c = rng.uniform(-1, 1, (8, 8)) * scale
        base = _ssd32(r, c)
        assert _ssd32(r[rows_eo], c[rows_eo]) == base
        for perm in cols.values():
            assert _ssd32(r[:, perm], c[:, perm]) == base
        p02 = [2, 1, 0, 3, 4, 5, 6, 7]
        changed_rows02 += _ssd32(r[p02], c[p02]) != base
        p12 = [0, 2, 1, 3, 4, 5, 6, 7]
        changed_cols12 += _ssd32(r[:, p12], c[:, p12]) != base
    assert changed_rows02 > 20 and changed_cols12 > 20


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


# ---------------------------------------------------- extractor pre-rounding
# The GPU kernel pre-rounds every weighted estimate and weight as
# v <- (v + E) - E with E = 2^extractor_exp (source/kernel.cu:740-742). The
# oracle's (v + E) - E is pinned against binary32 round-to-nearest-even
# evaluated in exact rational arithmetic, against a closed form on
# [-E/2, E), and through the grid every pre-rounded value lies on.
from fractions import Fraction


def _rne32(x: Fraction) -> Fraction:
    """x rounded to the nearest binary32 value, ties to even (exact)."""
    if x == 0:
        return Fraction(0)
    a = abs(x)
    k = a.numerator.bit_length() - a.denominator.bit_length()
    if Fraction(2) ** k > a:
        k -= 1
    grid = Fraction(2) ** (max(k, -126) - 23)
    q = a / grid
    n = q.numerator // q.denominator
    rem = q - n
    if rem > Fraction(1, 2) or (rem == Fraction(1, 2) and n % 2 == 1):
        n += 1
    return (n * grid) if x > 0 else -(n * grid)


def _preround_samples(e, rng):
    E = 2.0 ** e
    g = 2.0 ** (e - 24)
    v = [rng.uniform(-E, E, 300), rng.uniform(-E / 2, E / 2, 200), rng.uniform(-8 * E, 8 * E, 200),
         rng.normal(size=100) * 2.0 ** rng.integers(-40, 0, 100),      # tiny values
         (rng.integers(-2 ** 23, 2 ** 23, 100) + 0.5) * g,               # ties on the 2^(e-24) grid
         (rng.integers(-2 ** 22, 2 ** 22, 100) + 0.5) * 2 * g,           # ties on the 2^(e-23) grid
         -E / 2 + rng.integers(-50, 50, 100) * g / 4,                     # around -E/2
         np.array([0.0, -0.0, E, -E, E / 2, -E / 2, 2 * E, -2 * E])]
    return np.concatenate(v).astype(np.float32)


@pytest.mark.parametrize("e", [0, 3, 6, 10, 20])
def test_preround_is_binary32_rne(e):
    """(v + E) - E == rne32(rne32(v + E) - E), computed exactly."""
    E = np.float32(2.0 ** e)
    rng = np.random.default_rng(100 + e)
    for v in _preround_samples(e, rng):
        want = _rne32(_rne32(Fraction(float(v)) + Fraction(float(E))) - Fraction(float(E)))
        got = O.preround(v, E)
        assert Fraction(got) == want, (float(v), e, got, float(want))


@pytest.mark.parametrize("e", [0, 4, 9])
def test_preround_closed_form(e):
    """On [-E/2, E) the pre-rounding is v rounded half-to-even to the grid
    2^(e-24) (v < 0) or 2^(e-23) (v >= 0): the ulp of v + E, exact in the
    subtraction (Sterbenz)."""
    E = 2.0 ** e
    rng = np.random.default_rng(7 + e)
    v = np.concatenate([rng.uniform(-E / 2, E, 2000),
                        (rng.integers(-2 ** 22, 2 ** 23, 300) + 0.5) * 2.0 ** (e - 24)]).astype(np.float32)
    v = v[(v >= -E / 2) & (v < E)].astype(np.float64)
    g = np.where(v >= 0, 2.0 ** (e - 23), 2.0 ** (e - 24))
    want = np.rint(v / g) * g   # exact: power-of-two scaling, rint ties to even
    got = np.array([O.preround(x, E) for x in v])
    np.testing.assert_array_equal(got, want)


@pytest.mark.parametrize("e", [0, 5, 12])
def test_preround_values_on_unit_grid(e):
    """Every pre-rounded value is a whole multiple of 2^(e-24), the unit of
    the kernels' fixed-point accumulation."""
    rng = np.random.default_rng(31 + e)
    unit = 2.0 ** (e - 24)
    for v in _preround_samples(e, rng):
        q = O.preround(v, 2.0 ** e) / unit
        assert float(q).is_integer(), (float(v), e)


def _coverage(groups, bs, H, W, tw):
    """Members covering each pixel of each slot: counts [tw, H, W]."""
    cnt = np.zeros((tw, H, W), np.int64)
    for x, y, z in groups.reshape(-1, 3):
        cnt[z, y:y + 8, x:x + 8] += 1
    return cnt


@pytest.mark.parametrize("final", [False, True])
@pytest.mark.parametrize("e", [2, 8])
def test_extractor_stacks_on_grid_and_bounded(final, e):
    """With extractor_exp = e every num and den entry of an oracle stack is a
    multiple of 2^(e-24) (a dropped rounding of w or of w * est breaks it),
    and |stack_E - stack_0| <= contributions * 2^(e-24) plus float
    accumulation slack (each contribution moves by at most half its grid)."""
    H, W, N, r = 32, 40, 3, 1
    bs, R, pn, pr = 4, 4, 2, 2
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=15, seed=5, motion=True)
    clip = noisy[:, 0]
    win = np.stack([clip[min(max(1 - r + z, 0), N - 1)] for z in range(2 * r + 1)])[None]
    refw = None
    if final:
        base = O.bm3d_window(win, None, [O.scale_sigma(15.0, False)], bs, R, r, pn, pr, False, False)
        refw = np.where(base[:, :, 1] > 0, base[:, :, 0] / np.maximum(base[:, :, 1], 1e-30), win)
    sig = [O.scale_sigma(15.0, final)]
    E = float(2.0 ** e)
    sE, groups, _ = O.bm3d_window(win, refw, sig, bs, R, r, pn, pr, False, final, E, want_groups=True)
    s0 = O.bm3d_window(win, refw, sig, bs, R, r, pn, pr, False, final, 0.0)
    unit = 2.0 ** (e - 24)
    q = sE.astype(np.float64) / unit
    assert np.array_equal(q, np.rint(q)), "stack entry off the 2^(e-24) grid"
    # control: without pre-rounding the weights are not all on that grid
    q0 = s0[:, :, 1].astype(np.float64) / unit
    assert not np.array_equal(q0, np.rint(q0))
    cnt = _coverage(groups, bs, H, W, 2 * r + 1)[None]
    for part in (0, 1):
        a = sE[:, :, part].astype(np.float64)
        b = s0[:, :, part].astype(np.float64)
        slack = cnt * 2.0 ** -23 * np.maximum(np.abs(a), np.abs(b))
        assert np.all(np.abs(a - b) <= cnt * unit + slack)
    if e >= 8:
        assert np.abs(sE - s0).max() > 0  # the rounding is visible at this E
