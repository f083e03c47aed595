"""The mixed matcher's spatial search on the matrix cores (the opt-in
matching kernel 2; the default scan runs on the vector ALUs).

Where a work-group tile allows (16 blocks per row, all block_step apart
inside the plane, block_step dividing 8, a union row of at most 80
positions), the spatial search's cross terms run as FP16 matrix products.
Its pre-scores differ from the vector-ALU scan's only in FP32 summation
order, and both are followed by the same exact rescoring, so:
  * the group lists equal the FP32 matcher's except among exactly tied
    candidates (basic pass, noisy clips) or among candidates within 2^-7 of
    each other (final pass on a smooth, partly clipped estimate, where many
    distances nearly tie), with the matrix-core search on ("mixed-mfma") and off
    ("mixed", the default), on clips whose edge tiles take the vector-ALU path;
  * the in-plane candidate counts equal the FP32 matcher's (whenever the
    lists do: the temporal windows are centred on them);
  * with extractor_exp the stacks of both forms are bit-identical whenever
    their lists are.
"""
import numpy as np
import pytest

import mixed_ties
from bm3d_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _core():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import core
    return core


# (H, W, N, block_step, bm_range, radius, ps_num, ps_range)
CASES = [
    (200, 264, 5, 4, 9, 1, 2, 4),   # the benchmark's parameters; the last tile column is irregular
    (192, 256, 3, 4, 9, 0, 2, 4),   # radius 0, every tile regular
    (150, 136, 4, 2, 7, 1, 2, 3),   # block_step 2
    (96, 120, 3, 1, 5, 2, 3, 2),    # block_step 1, radius 2
    (180, 232, 3, 4, 5, 1, 2, 4),   # a short window
    (99, 147, 3, 4, 9, 1, 2, 4),    # odd sizes: clamped last row and column
]


@pytest.mark.parametrize("case", CASES, ids=[f"{h}x{w}-bs{b}-R{R}-r{r}" for h, w, _, b, R, r, _, _ in CASES])
@pytest.mark.parametrize("final", [False, True], ids=["basic", "final"])
def test_lists_counts_and_stacks(case, final):
    core = _core()
    H, W, N, bs, R, r, pn, pr = case
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=700 + bs + R + r, motion=True)
    src = torch.from_numpy(noisy).cuda()
    ref = torch.from_numpy(np.clip(clean + np.float32(0.002), 0, 1).astype(np.float32)).cuda() if final else None
    kw = dict(block_step=bs, bm_range=R, radius=r, ps_num=pn, ps_range=pr, final=final, extractor_exp=10)
    pm = core.Plan(W, H, 15.0, precision="mixed", **kw)
    pm.select_kernels("mixed-mfma", "mixed")
    pv = core.Plan(W, H, 15.0, precision="mixed", **kw)
    p32 = core.Plan(W, H, 15.0, **kw)
    assert pm.uses_matrix_core_search() and not pv.uses_matrix_core_search()
    gm = core.unpack_groups(pm.match(src, ref))
    gv = core.unpack_groups(pv.match(src, ref))
    g32 = core.unpack_groups(p32.match(src, ref))
    clip_m = noisy if ref is None else ref.cpu().numpy()
    rel = 2.0 ** -7 if final else 0.0
    for name, got in (("mixed-mfma", gm), ("mixed", gv)):
        nd, nl = mixed_ties.tie_only(got, g32, clip_m, r, bs, rel=rel)
        print(f"{name}: {nd} of {nl} lists differ from FP32, all at {'near ' if rel else 'exact '}ties")
    # in-plane candidates: fixed by the geometry at radius 0; with a temporal
    # search, by the lists that centre its windows (equal when the lists are)
    if r == 0 or np.array_equal(gm, g32):
        assert pm.count_candidates(src, ref) == p32.count_candidates(src, ref)
    sm, sv = pm.execute(src, ref), pv.execute(src, ref)
    same = np.all(gm == gv, axis=(2, 3))  # [frames, blocks]
    if same.all():
        assert torch.equal(sm, sv)
