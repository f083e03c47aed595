"""Batched multi-clip launches (bm3d_batch_*) give every clip the values of
its own plans: BM3Dv2 per pass through bm3d_execute_vagg, unprocessed planes
copied from the source (source/source.cpp:911-989). With extractor_exp set
the sums are exact, so the comparison is bit for bit -- across clip sizes,
radii 0..2, one or two passes, FP32 and mixed precision, chroma, several
waves (a tiny scratch budget) and repeated graph replays. The FP32 jobs
are also checked against the oracle's pipeline."""
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


JOBS = [
    # (H, W, N, planes, params)
    (48, 64, 5, 1, dict(sigma=[12.0], block_step=4, bm_range=5, radius=1, ps_num=2, ps_range=3, passes=2,
                        precision="fp32")),
    (40, 56, 4, 1, dict(sigma=[20.0], block_step=3, bm_range=4, radius=0, passes=2, precision="mixed")),
    (40, 48, 6, 3, dict(sigma=[10.0, 0.0, 12.0], block_step=4, bm_range=4, radius=2, ps_num=2, ps_range=2,
                        chroma=True, passes=1, precision="fp32")),
    (72, 88, 7, 1, dict(sigma=[15.0], block_step=5, bm_range=6, radius=2, ps_num=3, ps_range=2, passes=2,
                        precision="mixed")),
    (33, 47, 3, 1, dict(sigma=[25.0], block_step=8, bm_range=3, radius=1, ps_num=1, ps_range=1, passes=1,
                        precision="mixed")),
    (24, 40, 2, 1, dict(sigma=[0.0], block_step=4, bm_range=4, radius=1, passes=2, precision="fp32")),
]


def _make_jobs(extractor_exp=10):
    jobs = []
    for k, (H, W, N, P, prm) in enumerate(JOBS):
        chroma = prm.get("chroma", False)
        _, noisy = synth.make_clip(H, W, N, P, noise_sd=max(prm["sigma"][0], 5.0), seed=40 + k, motion=True,
                                   yuv=chroma)
        j = dict(prm)
        j["src"] = torch.from_numpy(noisy).cuda()
        j["extractor_exp"] = extractor_exp
        jobs.append(j)
    return jobs


def _expected(core, j):
    src = j["src"]
    n, P, H, W = src.shape
    kw = {k: j[k] for k in ("block_step", "bm_range", "radius", "ps_num", "ps_range", "chroma", "extractor_exp",
                            "precision") if k in j}
    sig = j["sigma"]
    process = [not (s < np.finfo(np.float32).eps) for s in (list(sig) + [sig[-1]] * 3)[:P]]

    def run(final, ref):
        if not any(process):
            return src.clone()
        plan = core.Plan(W, H, sig, final=final, **kw)
        out = plan.execute_vagg(src, ref)
        for p in range(P):
            if not process[p]:
                out[:, p] = src[:, p]
        return out

    basic = run(False, None)
    if j["passes"] == 1:
        return basic.cpu().numpy(), None
    return run(True, basic).cpu().numpy(), basic.cpu().numpy()


@pytest.mark.parametrize("budget", [0, 1])
def test_batch_equals_plans(budget):
    core = _core()
    jobs = _make_jobs()
    want = [_expected(core, j) for j in jobs]
    batch = core.Batch(jobs, scratch_budget=budget)
    info = batch.info
    assert info["waves"] == (1 if budget == 0 else len(JOBS))  # budget 1: a wave per job
    for rep in range(3):  # graph replays, with eager work between them
        for bj in batch.jobs:
            bj["out"].fill_(-7.0)
            if bj["basic"] is not None:
                bj["basic"].fill_(-7.0)
        batch.execute()
        torch.cuda.synchronize()
        for k, (bj, (w_out, w_basic)) in enumerate(zip(batch.jobs, want)):
            np.testing.assert_array_equal(bj["out"].cpu().numpy(), w_out, err_msg=f"job {k} replay {rep}")
            if w_basic is not None:
                np.testing.assert_array_equal(bj["basic"].cpu().numpy(), w_basic, err_msg=f"job {k} basic")


def test_batch_fp32_jobs_equal_oracle():
    """The FP32 jobs of the batch (two passes at radius 1; chroma at radius 2
    with an unprocessed plane; a clip passed through for sigma 0) against
    the oracle's BM3Dv2 pipeline, bit for bit (extractor_exp set)."""
    core = _core()
    jobs = _make_jobs()
    batch = core.Batch(jobs)
    batch.execute()
    torch.cuda.synchronize()
    checked = 0
    for k, (bj, (H, W, N, P, prm)) in enumerate(zip(batch.jobs, JOBS)):
        if prm["precision"] != "fp32":
            continue
        noisy = jobs[k]["src"].cpu().numpy()
        p = O.Params(sigma=tuple(prm["sigma"]), block_step=prm["block_step"], bm_range=prm["bm_range"],
                     radius=prm["radius"], ps_num=prm.get("ps_num", 2), ps_range=prm.get("ps_range", 4),
                     chroma=prm.get("chroma", False), extractor_exp=10)
        if prm["passes"] == 2:
            want_b, want = O.two_pass(noisy, p)
            np.testing.assert_array_equal(bj["basic"].cpu().numpy(), want_b, err_msg=f"job {k} basic")
        else:
            want = O.bm3dv2(noisy, None, p)
        np.testing.assert_array_equal(bj["out"].cpu().numpy(), want, err_msg=f"job {k}")
        checked += 1
    assert checked == 3


def test_batch_without_extractor_close_to_plans():
    core = _core()
    jobs = _make_jobs(extractor_exp=0)
    want = [_expected(core, j) for j in jobs]
    batch = core.Batch(jobs)
    batch.execute()
    for k, (bj, (w_out, _)) in enumerate(zip(batch.jobs, want)):
        np.testing.assert_allclose(bj["out"].cpu().numpy(), w_out, rtol=2e-5, atol=2e-5, err_msg=f"job {k}")


def test_batch_shapes_and_autotune_keep_results():
    core = _core()
    jobs = _make_jobs()
    batch = core.Batch(jobs)
    batch.execute()
    want = [bj["out"].cpu().numpy() for bj in batch.jobs]
    for t in [dict(match_rows=4), dict(match_width=32, match_rows=8),
              dict(filter_tile_x=16, filter_tile_y=8, filter_waves=8),
              dict(filter_tile_x=4, filter_tile_y=4, filter_waves=1)]:
        batch.set_tuning(**t)
        batch.execute()
        for k, bj in enumerate(batch.jobs):
            np.testing.assert_array_equal(bj["out"].cpu().numpy(), want[k], err_msg=f"{t} job {k}")
    batch.autotune()
    batch.execute()
    for k, bj in enumerate(batch.jobs):
        np.testing.assert_array_equal(bj["out"].cpu().numpy(), want[k], err_msg=f"autotuned job {k}")


def test_batch_rejects():
    core = _core()
    from bm3d_amd import _capi as C
    jobs = _make_jobs()[:1]
    jobs[0]["passes"] = 3
    with pytest.raises(C.BM3DError, match="passes"):
        core.Batch(jobs)
    jobs[0]["passes"] = 1
    jobs[0]["block_step"] = 9
    with pytest.raises(C.BM3DError, match="block_step"):
        core.Batch(jobs)


def test_execute_vagg_radius0_equals_aggregate():
    """bm3d_execute_vagg with radius 0 is BM3D's aggregated output
    (bm3d_execute + bm3d_aggregate, source/source.cpp:170-188)."""
    core = _core()
    _, noisy = synth.make_clip(48, 64, 3, 1, noise_sd=20.0, seed=47)
    src = torch.from_numpy(noisy).cuda()
    for prec in ("fp32", "mixed"):
        plan = core.Plan(64, 48, 20.0, block_step=3, bm_range=5, radius=0, extractor_exp=9, precision=prec)
        want = core.aggregate(plan.execute(src)).cpu().numpy()
        np.testing.assert_array_equal(plan.execute_vagg(src).cpu().numpy(), want)
        np.testing.assert_array_equal(plan.execute_vagg(src, first=1, count=2).cpu().numpy(), want[1:])
