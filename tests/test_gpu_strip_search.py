"""The mixed matcher's strip search (matching kernel 3, the default at block
steps 2 and 4).

Reference blocks bs apart overlap, so a block's 8 columns are the strips (bs
columns x 8 rows) of 8 / bs neighbouring lanes: each lane pre-scores its own
strip at every displacement and adds its neighbours' strips by DPP. The
pre-scores differ from the per-lane scan's (kernel 1) in FP32 summation
order only, and the same exact rescoring follows, so:
  * the group lists equal the FP32 matcher's except among exactly tied
    candidates (basic pass) or near-tied ones, within 2^-7 (final pass on a
    smooth estimate), with the strip search on and off; tiles whose last block
    is clamped to W - 8 take the per-lane scan;
  * the in-plane candidate counts equal the FP32 matcher's whenever the lists
    do (the temporal windows are centred on them);
  * with extractor_exp the stacks of both searches are bit-identical wherever
    their lists are.
Rows of 16 lanes (bs 2 and 4) and of 32 (bs 4) are covered; bs 2 with rows of
32, and other block steps, keep the per-lane scan. At radius > 0 a plan runs
the strip search as two launches (spatial by strips, then temporal with one
lane per block); bm3d_execute_noalloc runs it in one launch, and
test_zz_no_alloc.py holds the two forms' stacks bit-identical.
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


# (H, W, N, block_step, bm_range, radius, ps_num, ps_range, match_width)
CASES = [
    (200, 264, 5, 4, 9, 1, 2, 4, 16),   # the benchmark's parameters
    (200, 264, 5, 4, 9, 1, 2, 4, 32),   # rows of 32 lanes (wave-wide DPP)
    (192, 256, 3, 4, 9, 0, 2, 4, 16),   # radius 0
    (150, 136, 4, 2, 7, 1, 2, 3, 16),   # block_step 2
    (120, 200, 3, 2, 12, 2, 3, 5, 16),  # block_step 2, radius 2
    (180, 232, 3, 4, 5, 1, 2, 4, 32),   # a short window
    (99, 147, 3, 4, 9, 1, 2, 4, 16),    # odd sizes: last row and column clamped
    (40, 48, 3, 4, 9, 1, 2, 4, 16),     # every window cut by the plane
]


def _ids():
    return [f"{h}x{w}-bs{b}-R{R}-r{r}-w{mw}" for h, w, _, b, R, r, _, _, mw in CASES]


@pytest.mark.parametrize("case", CASES, ids=_ids())
@pytest.mark.parametrize("final", [False, True], ids=["basic", "final"])
def test_strip_lists_counts_and_stacks(case, final):
    core = _core()
    H, W, N, bs, R, r, pn, pr, mw = case
    clean, noisy = synth.make_clip(H, W, N, 1, noise_sd=15.0, seed=800 + bs + R + r, motion=True)
    src = torch.from_numpy(noisy).cuda()
    ref = torch.from_numpy(np.clip(clean + np.float32(0.002), 0, 1).astype(np.float32)).cuda() if final else None
    kw = dict(block_step=bs, bm_range=R, radius=r, ps_num=pn, ps_range=pr, final=final, extractor_exp=10)
    ps = core.Plan(W, H, 15.0, precision="mixed", **kw)
    pl = core.Plan(W, H, 15.0, precision="mixed", **kw)
    pl.select_kernels("mixed", "mixed")
    p32 = core.Plan(W, H, 15.0, **kw)
    for p in (ps, pl):
        p.set_tuning(match_width=mw, match_rows=8 if mw == 32 else 16)
    assert ps.uses_strip_search() and not pl.uses_strip_search() and pl.uses_mixed_kernel()
    gs = core.unpack_groups(ps.match(src, ref))
    gl = core.unpack_groups(pl.match(src, ref))
    g32 = core.unpack_groups(p32.match(src, ref))
    clip_m = noisy if ref is None else ref.cpu().numpy()
    rel = 2.0 ** -7 if final else 0.0
    for name, got in (("strip", gs), ("per-lane", gl)):
        nd, nl = mixed_ties.tie_only(got, g32, clip_m, r, bs, rel=rel)
        print(f"{name}: {nd} of {nl} lists differ from FP32, all at {'near ' if rel else 'exact '}ties")
    if r == 0 or np.array_equal(gs, g32):
        assert ps.count_candidates(src, ref) == p32.count_candidates(src, ref)
    ss, sl = ps.execute(src, ref), pl.execute(src, ref)
    same = np.all(gs == gl, axis=(2, 3))  # [frames, blocks]
    if same.all():
        assert torch.equal(ss, sl)


def test_strip_search_selection():
    """Strips serve block steps 4 (rows of 16 or 32 lanes) and 2 (rows of
    16), by default; elsewhere a plan keeps the per-lane search, and selecting
    the strip search there is not an error."""
    core = _core()
    W, H = 128, 96
    for bs, mw, want in [(4, 16, True), (4, 32, True), (2, 16, True), (2, 32, False), (8, 16, False),
                         (3, 16, False), (1, 16, False)]:
        p = core.Plan(W, H, 10.0, block_step=bs, bm_range=6, radius=0, precision="mixed")
        p.set_tuning(match_width=mw, match_rows=8)
        assert p.uses_mixed_kernel()
        assert p.uses_strip_search() == want, (bs, mw)
        p1 = core.Plan(W, H, 10.0, block_step=bs, bm_range=6, radius=1, precision="mixed")
        p1.set_tuning(match_width=mw, match_rows=8)
        assert p1.uses_strip_search() == want
        p.select_kernels("mixed", "mixed")
        assert not p.uses_strip_search()
        p.select_kernels("mixed-strip", "mixed")
        assert p.uses_strip_search() == want
        assert p.uses_mixed_kernel()

