"""Group-list contract of the mixed-precision matcher: its lists equal the
FP32 matcher's except where candidates tie exactly. A differing list passes
only when the two lists hold the same multiset of exact FP32 distances
(the reference's summation order, recomputed by the oracle), i.e. they
differ only in which of several equally distant candidates they keep."""
import numpy as np

from oracle import oracle as O


def tie_only(gm, g32, clip, radius, block_step, first=0, rel=0.0):
    """gm, g32: unpacked lists [frames, blocks, 8, 3] (x, y, z) of output
    frames first.. of `clip` [N, planes, H, W] (plane 0 is matched).
    Returns (number of differing lists, number of lists); asserts that every
    difference is among tied candidates -- or, with rel > 0, among candidates
    whose exact distances lie within a relative `rel` of each other (the
    FP32 list's sorted distances are the smallest possible, entry by entry;
    the mixed list's may exceed them by at most rel times)."""
    N, _, H, W = clip.shape
    xs = O.block_positions(W, block_step)
    ys = O.block_positions(H, block_step)
    diff = np.nonzero(np.any(gm != g32, axis=(2, 3)))
    for fi, blk in zip(*diff):
        n = first + int(fi)
        by, bx = divmod(int(blk), len(xs))
        x, y = xs[bx], ys[by]
        ref = np.ascontiguousarray(clip[n, 0, y:y + 8, x:x + 8])

        def dists(lst):
            out = []
            for px, py, z in lst:
                f = min(max(n - radius + int(z), 0), N - 1)
                out.append(O.ssd(ref, clip[f, 0], int(px), int(py)))
            return sorted(out)

        a, b = dists(gm[fi, blk]), dists(g32[fi, blk])
        if rel == 0.0:
            assert a == b, f"frame {n} block ({x}, {y}): mixed {a} vs fp32 {b}"
        else:
            assert all(bi <= ai <= bi * (1.0 + rel) for ai, bi in zip(a, b)), \
                f"frame {n} block ({x}, {y}): mixed {a} vs fp32 {b}"
    return len(diff[0]), gm.shape[0] * gm.shape[1]
