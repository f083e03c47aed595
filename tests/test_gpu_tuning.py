"""Launch shapes (bm3d_plan_set_tuning / bm3d_plan_autotune) change speed,
never results.

The fused BM3D + VAggregate path accumulates, per output frame, the slots
of every evaluation VAggregate reads for it (source/source.cpp:816-846,
with the slot multiplicities of the clip's ends). Small batches and several
evaluation runs per launch make output frames span launches and runs; those
regions are summed atomically. With extractor_exp set every sum is exact, so
all shapes must give bm3d_execute + bm3d_vaggregate's values bit for bit.
"""
import numpy as np
import pytest

from bm3d_amd import synth

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


def _core():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import core
    return core


CASES = {
    "r1": (48, 64, 7, 1, False, dict(sigma=[12.0], block_step=4, bm_range=5, radius=1, ps_num=2, ps_range=3)),
    "r2": (40, 56, 9, 1, False, dict(sigma=[15.0], block_step=3, bm_range=4, radius=2, ps_num=2, ps_range=2)),
    "chroma-r1": (40, 48, 6, 3, True, dict(sigma=[10.0, 0.0, 12.0], block_step=4, bm_range=4, radius=1,
                                           ps_num=2, ps_range=2)),
}

TUNINGS = [
    dict(batch_frames=1, chunks=1),
    dict(batch_frames=2, chunks=1),
    dict(batch_frames=4, chunks=2),
    dict(batch_frames=5, chunks=5),
    dict(batch_frames=64, chunks=3),
    dict(filter_tile_x=16, filter_tile_y=8, filter_waves=8),
    dict(filter_tile_x=4, filter_tile_y=4, filter_waves=2, batch_frames=3, chunks=2),
    dict(filter_tile_x=3, filter_tile_y=5, filter_waves=1),
    dict(match_rows=4),
    dict(match_rows=8, filter_tile_x=16, filter_tile_y=16),
    dict(match_width=32, match_rows=8),
    dict(match_width=32, match_rows=4, filter_tile_x=16, filter_tile_y=8),
]


def _plan(core, W, H, prm, precision, final=False):
    return core.Plan(W, H, prm["sigma"], prm["block_step"], prm["bm_range"], prm["radius"], prm["ps_num"],
                     prm["ps_range"], chroma=len(prm["sigma"]) == 3, final=final, extractor_exp=10,
                     precision=precision)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("name", list(CASES))
def test_fused_shapes_equal_stacks(name, precision):
    core = _core()
    H, W, N, P, chroma, prm = CASES[name]
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=21, motion=True, yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    keep = [p for p in range(P) if not (chroma and prm["sigma"][p] == 0.0)]
    ref_plan = _plan(core, W, H, prm, precision)
    want = core.vaggregate(ref_plan.execute(src), prm["radius"], N).cpu().numpy()[:, keep]
    # final pass too, with the basic estimate as ref
    basic = torch.from_numpy(np.ascontiguousarray(want if P == 1 else noisy)).cuda()
    fin_plan = _plan(core, W, H, prm, precision, final=True)
    want_f = core.vaggregate(fin_plan.execute(src, basic), prm["radius"], N).cpu().numpy()[:, keep]
    for t in TUNINGS:
        plan = _plan(core, W, H, prm, precision)
        plan.set_tuning(**t)
        for first, count in [(0, N), (1, N - 2), (N - 1, 1)]:
            got = plan.execute_vagg(src, first=first, count=count).cpu().numpy()[:, keep]
            np.testing.assert_array_equal(got, want[first:first + count], err_msg=f"{t} {first}+{count}")
        plan_f = _plan(core, W, H, prm, precision, final=True)
        plan_f.set_tuning(**t)
        got_f = plan_f.execute_vagg(src, basic).cpu().numpy()[:, keep]
        np.testing.assert_array_equal(got_f, want_f, err_msg=f"final {t}")


@pytest.mark.parametrize("name", ["r1", "chroma-r1"])
def test_stack_shapes_equal(name):
    """Slot-mode filtering (bm3d_execute) with other tiles and batches."""
    core = _core()
    H, W, N, P, chroma, prm = CASES[name]
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=prm["sigma"][0], seed=22, motion=True, yuv=chroma)
    src = torch.from_numpy(noisy).cuda()
    want = _plan(core, W, H, prm, "fp32").execute(src).cpu().numpy()
    for t in TUNINGS:
        plan = _plan(core, W, H, prm, "fp32")
        plan.set_tuning(**t)
        np.testing.assert_array_equal(plan.execute(src).cpu().numpy(), want, err_msg=str(t))


@pytest.mark.parametrize("bad,msg", [
    (dict(match_rows=5), "match_rows"),
    (dict(match_width=8), "match_width"),
    (dict(match_width=32, match_rows=16), "at most 256"),
    (dict(filter_tile_x=0), "filter tile"),
    (dict(filter_tile_x=32, filter_tile_y=16), "filter tile"),
    (dict(filter_waves=9), "filter_waves"),
    (dict(batch_frames=0), "batch_frames"),
    (dict(batch_frames=4, chunks=5), "chunks"),
])
def test_set_tuning_rejects(bad, msg):
    core = _core()
    from bm3d_amd import _capi as C
    H, W, N, P, chroma, prm = CASES["r1"]
    plan = _plan(core, W, H, prm, "fp32")
    before = plan.tuning
    with pytest.raises(C.BM3DError) as ei:
        plan.set_tuning(**bad)
    assert ei.value.code == C.BM3D_EINVAL and msg in str(ei.value)
    assert plan.tuning == before


def test_autotune_keeps_results():
    core = _core()
    H, W, N, P, chroma, prm = CASES["r2"]
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=15.0, seed=23, motion=True)
    src = torch.from_numpy(noisy).cuda()
    plan = _plan(core, W, H, prm, "mixed")
    want = plan.execute_vagg(src).cpu().numpy()
    t = plan.autotune(src)
    assert set(t) == {"match_rows", "filter_tile_x", "filter_tile_y", "filter_waves", "batch_frames", "chunks",
                      "match_width"}
    np.testing.assert_array_equal(plan.execute_vagg(src).cpu().numpy(), want)
    # a new plan with the same parameters starts from the tuned shapes
    assert _plan(core, W, H, prm, "mixed").tuning == t
