"""Pins of the oracle's NVRTC-variant options (rtc_source/kernel.hpp:103-504):
the block distances ssd / sad / zssd / zsad / ssd/norm and the transforms
dct / haar / wht / bior1.5, each checked against an independent definition:

* Haar: the orthonormal Haar basis (scaling function, wavelets of widths
  8, 4, 2 in the reference's output order, kernel.hpp:225-232) times 4. The
  reference's inverse (kernel.hpp:239-259) reads coefficient 2 for 3 and
  coefficient 4 for 5, 6 and 7; it is pinned as the transpose of the forward
  matrix composed with that substitution.
* WHT: the Sylvester Hadamard matrix with its rows sorted by sign changes
  (sequency order), times sqrt(2); self-inverse up to 16.
* bior1.5: inverse(forward(x)) = 16 x, row 0 = the scaled mean, the finest
  level = the Haar pair differences, vanishing moments of the detail rows.
* distances: float64 brute force of each formula, exact invariants on
  integer-valued blocks (zssd / zsad ignore a constant offset; sad is
  symmetric), ssd/norm limits (c vs c -> 0, c vs -c -> 4, scale invariance).
* selection: the oracle's matcher with each distance against an order-free
  sort of every candidate by (distance, arrival index).
"""
import numpy as np
import pytest

from oracle import oracle as O

S2 = np.sqrt(2.0)


def haar_forward():
    rows = [np.full(8, 1 / np.sqrt(8)), np.r_[np.ones(4), -np.ones(4)] / np.sqrt(8)]
    for lo in (0, 4):
        r = np.zeros(8)
        r[lo:lo + 2], r[lo + 2:lo + 4] = 0.5, -0.5
        rows.append(r)
    for lo in (0, 2, 4, 6):
        r = np.zeros(8)
        r[lo], r[lo + 1] = 1 / S2, -1 / S2
        rows.append(r)
    return 4 * np.array(rows)


def haar_reference_inverse():
    """Transpose of the forward matrix, fed coefficient 2 in place of 3 and
    coefficient 4 in place of 5, 6, 7 (what kernel.hpp:239-259 computes)."""
    S = np.zeros((8, 8))
    for m, k in [(0, 0), (1, 1), (2, 2), (3, 2), (4, 4), (5, 4), (6, 4), (7, 4)]:
        S[m, k] = 1
    return haar_forward().T @ S


def walsh():
    h = np.array([[1.0]])
    for _ in range(3):
        h = np.block([[h, h], [h, -h]])
    changes = [(np.diff(np.sign(r)) != 0).sum() for r in h]
    return S2 * h[np.argsort(changes)]


# ------------------------------------------------------------- transforms
def test_haar_forward_is_4x_orthonormal_haar():
    np.testing.assert_allclose(O.xform_matrix(1), haar_forward(), atol=1e-15)
    H = O.xform_matrix(1)
    np.testing.assert_allclose(H @ H.T, 16 * np.eye(8), atol=1e-12)


def test_haar_inverse_is_the_references_map():
    Hi = O.xform_matrix(1, inverse=True)
    np.testing.assert_allclose(Hi, haar_reference_inverse(), atol=1e-15)
    # coefficients 3, 5, 6, 7 are never read, so it is not an inverse
    assert np.all(Hi[:, [3, 5, 6, 7]] == 0)
    assert not np.allclose(Hi @ O.xform_matrix(1), 16 * np.eye(8))
    # ... but it inverts the forward transform on signals whose coefficient 3
    # equals coefficient 2 and 5, 6, 7 equal coefficient 4
    c = np.array([1.0, -2.0, 0.5, 0.5, 3.0, 3.0, 3.0, 3.0])
    x = O.xform_matrix(1).T @ c / 16
    np.testing.assert_allclose(Hi @ (O.xform_matrix(1) @ x), 16 * x, atol=1e-12)


def test_wht_is_sequency_ordered_hadamard():
    W = walsh()
    np.testing.assert_allclose(O.xform_matrix(2), W, atol=1e-15)
    np.testing.assert_allclose(O.xform_matrix(2, inverse=True), W, atol=1e-15)
    np.testing.assert_allclose(W @ W, 16 * np.eye(8), atol=1e-12)


def test_bior15_pair():
    F, B = O.xform_matrix(3), O.xform_matrix(3, inverse=True)
    np.testing.assert_allclose(B @ F, 16 * np.eye(8), atol=1e-12)
    np.testing.assert_allclose(F[0], np.full(8, S2), atol=1e-15)
    np.testing.assert_allclose(F[4:], haar_forward()[4:], atol=1e-15)
    # detail rows annihilate constants; the level-2 details also annihilate ramps
    np.testing.assert_allclose(F[1:] @ np.ones(8), 0, atol=1e-12)
    ramp = np.arange(4.0)
    np.testing.assert_allclose(F[2, :4] @ ramp + F[2, 4:] @ ramp, F[2] @ np.r_[ramp, ramp], atol=1e-12)
    assert not np.allclose(F @ F.T, 16 * np.eye(8))  # biorthogonal, not orthogonal


def test_dct_matrix_matches_butterfly():
    x = np.random.default_rng(1).normal(size=(30, 8))
    np.testing.assert_allclose(O.xform8(x, 0, dtype=np.float64), x @ O.xform_matrix(0).T, atol=1e-12)
    np.testing.assert_allclose(O.xform8(x, 0, True, dtype=np.float64), x @ O.xform_matrix(0, True).T, atol=1e-12)


@pytest.mark.parametrize("kind", [1, 2, 3])
@pytest.mark.parametrize("inverse", [False, True])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 3e-6)])
def test_xform8_is_the_matrix(kind, inverse, dtype, tol):
    x = np.random.default_rng(2 + kind).normal(size=(40, 8))
    M = (haar_reference_inverse() if inverse else haar_forward()) if kind == 1 else \
        walsh() if kind == 2 else O.xform_matrix(3, inverse)
    np.testing.assert_allclose(O.xform8(x, kind, inverse, dtype=dtype), x @ M.T, rtol=0, atol=tol * 16)


@pytest.mark.parametrize("t2,t1", [(1, 0), (0, 2), (3, 3), (2, 1), (3, 0)])
def test_transform3d_kinds_separable(t2, t1):
    g = np.random.default_rng(7).normal(size=(8, 8, 8))
    A, B = O.xform_matrix(t2), O.xform_matrix(t1)
    want = np.einsum("ai,bj,ck,ijk->abc", B, A, A, g)
    got = O.transform3d(g, dtype=np.float64, t2=t2, t1=t1)
    np.testing.assert_allclose(got, want, atol=1e-10)
    Ai, Bi = O.xform_matrix(t2, True), O.xform_matrix(t1, True)
    back = O.transform3d(got, inverse=True, dtype=np.float64, t2=t2, t1=t1)
    np.testing.assert_allclose(back, np.einsum("ai,bj,ck,ijk->abc", Bi, Ai, Ai, got), atol=1e-8)
    if 1 not in (t2, t1):
        np.testing.assert_allclose(back, 4096 * g, atol=1e-8)


@pytest.mark.parametrize("t2,t1", [(2, 2), (3, 3), (2, 3), (0, 3)])
def test_collab_hard_sigma_zero_identity_other_transforms(t2, t1):
    """sigma = 0 keeps every coefficient: the group comes back unchanged
    (up to rounding) for every invertible pair, weight 1 / 512."""
    g = np.random.default_rng(8).uniform(size=(8, 8, 8))
    c = O.transform3d(g, dtype=np.float64, t2=t2, t1=t1) / 4096
    back = O.transform3d(c, inverse=True, dtype=np.float64, t2=t2, t1=t1)
    np.testing.assert_allclose(back, g, atol=1e-12)


# ------------------------------------------------------------- distances
def brute(cost, c, n):
    c = c.astype(np.float64)
    n = n.astype(np.float64)
    d = c - n
    if cost == 0:
        return np.sum(d * d)
    if cost == 1:
        return np.sum(np.abs(d))
    if cost in (2, 3):
        z = d - (c.mean() - n.mean())
        return np.sum(z * z) if cost == 2 else np.sum(np.abs(z))
    eps = float(np.finfo(np.float32).eps)
    return 2 - 2 * np.sum(c * n) / (np.sqrt(np.sum(c * c)) * np.sqrt(np.sum(n * n)) + eps)


@pytest.mark.parametrize("cost", [0, 1, 2, 3, 4])
@pytest.mark.parametrize("dtype,tol", [(np.float64, 1e-12), (np.float32, 2e-6)])
def test_cost_matches_float64(cost, dtype, tol):
    rng = np.random.default_rng(10 + cost)
    img = rng.uniform(size=(20, 24)).astype(dtype)
    ref = img[3:11, 5:13].copy()
    for (x, y) in [(0, 0), (5, 3), (16, 12), (9, 7), (2, 11)]:
        want = brute(cost, ref, img[y:y + 8, x:x + 8])
        got = O.block_cost(ref, img, x, y, cost, dtype=dtype)
        assert abs(got - want) <= tol * max(abs(want), 1.0), (cost, x, y, got, want)


def test_cost_ssd_is_the_ssd_path():
    rng = np.random.default_rng(15)
    img = rng.uniform(size=(16, 16)).astype(np.float32)
    ref = img[4:12, 2:10].copy()
    for (x, y) in [(0, 0), (3, 5), (8, 8)]:
        assert O.block_cost(ref, img, x, y, 0) == O.ssd(ref, img, x, y)


def test_cost_invariants_integer_blocks():
    rng = np.random.default_rng(16)
    c = rng.integers(-50, 50, size=(8, 8)).astype(np.float32)
    n = rng.integers(-50, 50, size=(8, 8)).astype(np.float32)
    for k in (1.0, -7.0, 64.0):
        # a constant offset leaves the zero-mean distances unchanged, exactly
        assert O.block_cost(c, n + np.float32(k), 0, 0, 2) == O.block_cost(c, n, 0, 0, 2)
        assert O.block_cost(c, n + np.float32(k), 0, 0, 3) == O.block_cost(c, n, 0, 0, 3)
    assert O.block_cost(c, n, 0, 0, 1) == O.block_cost(n, c, 0, 0, 1)
    assert O.block_cost(c, c, 0, 0, 1) == 0 and O.block_cost(c, c, 0, 0, 2) == 0
    # the zero-mean distance of c vs c + k is 0 while the plain one is not
    assert O.block_cost(c, c + np.float32(3), 0, 0, 3) == 0
    assert O.block_cost(c, c + np.float32(3), 0, 0, 1) == 64 * 3


def test_ssd_norm_limits():
    rng = np.random.default_rng(17)
    c = rng.uniform(0.1, 1, size=(8, 8)).astype(np.float32)
    assert abs(O.block_cost(c, c, 0, 0, 4)) < 1e-6
    assert abs(O.block_cost(c, -c, 0, 0, 4) - 4) < 1e-6
    n = rng.uniform(0.1, 1, size=(8, 8)).astype(np.float32)
    a = O.block_cost(c, n, 0, 0, 4)
    assert abs(O.block_cost(c, np.float32(4) * n, 0, 0, 4) - a) < 1e-6
    z = np.zeros((8, 8), np.float32)
    assert O.block_cost(c, z, 0, 0, 4) == 2.0


# ------------------------------------------------- selection per distance
def brute_lists(slots, bs, R, radius, ps_num, ps_range, cost):
    """Every list is the 8 smallest (distance, arrival index) keys seen."""
    H, W = slots[0].shape
    BIG = float(np.finfo(np.float32).max)
    out = []
    for y in O.block_positions(H, bs):
        for x in O.block_positions(W, bs):
            ref = slots[radius][y:y + 8, x:x + 8]
            arrival = [0]

            def window(cx, cy, rg, z):
                keys = []
                for row in range(max(cy - rg, 0), min(cy + rg, H - 8) + 1):
                    for col in range(max(cx - rg, 0), min(cx + rg, W - 8) + 1):
                        keys.append((O.block_cost(ref, slots[z], col, row, cost), arrival[0], col, row, z))
                        arrival[0] += 1
                return keys

            def top8(keys):
                return sorted(keys, key=lambda k: (k[0], k[1]))[:8]

            L = top8([(BIG, -1, x, y, radius)] * 8 + window(x, y, R, radius))
            if radius:
                centre = [(k[2], k[3]) for k in L]
                for d in (-1, 1):
                    last = centre
                    for t in range(1, radius + 1):
                        z = radius + d * t
                        F = [(BIG, -1, 0, 0, z)] * 8
                        for i in range(ps_num):
                            F = top8(F + window(last[i][0], last[i][1], ps_range, z))
                        L = top8(L + [(k[0], arrival[0] + i, k[2], k[3], k[4]) for i, k in enumerate(F[:ps_num])])
                        arrival[0] += ps_num
                        last = [(k[2], k[3]) for k in F]
            pos = [(k[2], k[3], k[4]) for k in L]
            if (x, y, radius) not in pos:
                pos = [(x, y, radius)] + pos[:7]
            out.append(pos)
    return np.array(out)


@pytest.mark.parametrize("cost", [1, 2, 3, 4])
@pytest.mark.parametrize("radius", [0, 1])
def test_selection_order_free(cost, radius):
    rng = np.random.default_rng(20 + cost + 5 * radius)
    base = rng.uniform(size=(18, 20)).astype(np.float32)
    slots = [np.ascontiguousarray(np.roll(base, (k, -k), axis=(0, 1)) + rng.normal(0, 0.05, base.shape),
                                  dtype=np.float32) for k in range(2 * radius + 1)]
    g, _, _ = O.match(slots, bs=4, bm_range=3, radius=radius, ps_num=2, ps_range=2, cost=cost)
    np.testing.assert_array_equal(g, brute_lists(slots, 4, 3, radius, 2, 2, cost))


def test_costs_select_differently():
    """A brightness-shifted copy: zssd finds it, ssd does not."""
    rng = np.random.default_rng(30)
    img = rng.uniform(size=(24, 24)).astype(np.float32)
    img[12:20, 12:20] = img[0:8, 0:8] + np.float32(0.5)
    g0, _, _ = O.match([img], bs=8, bm_range=12, cost=0)
    g2, _, _ = O.match([img], bs=8, bm_range=12, cost=2)
    blk = O.block_positions(24, 8).index(0) * len(O.block_positions(24, 8)) + O.block_positions(24, 8).index(0)
    assert (12, 12, 0) in [tuple(p) for p in g2[blk][:2]]
    assert (12, 12, 0) not in [tuple(p) for p in g0[blk][:2]]


@pytest.mark.parametrize("cost,t2,t1", [(1, 2, 2), (2, 3, 0), (4, 0, 3), (3, 1, 1)])
def test_pipeline_fp32_close_to_fp64(cost, t2, t1):
    from bm3d_amd import synth
    _, noisy = synth.make_clip(32, 40, 3, 1, noise_sd=10.0, seed=31, motion=True)
    p = O.Params(sigma=(10.0,), block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2,
                 bm_error=cost, transform_2d=t2, transform_1d=t1)
    a = O.bm3dv2(noisy, None, p, np.float32)
    b = O.bm3dv2(noisy.astype(np.float64), None, p, np.float64)
    ok = np.abs(a - b) < 1e-3 * max(1.0, np.abs(b).max())
    assert ok.mean() > 0.995
