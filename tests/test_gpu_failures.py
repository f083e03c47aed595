"""Runtime failures forced with bm3d_debug_inject_failure.

The reference maps every runtime failure to an error on the VapourSynth map
or frame context (setError / setFilterError, source/source.cpp:54-59,
225-231, 305-314). Here, for each kind of runtime event -- device
allocation, pinned host allocation, kernel launch -- and each position k in
a call's sequence of such events, the k-th event is made to fail. Bar:
  * the call reports BM3D_EHIP with a message naming the failed operation
    and the runtime's error string;
  * the same objects (plans, filters, batch, plugin instances) then
    produce results bit-identical to an undisturbed run (extractor_exp set);
  * sweeping k until a call completes without reaching the countdown
    visits every event of the call.
Covered: plans (two passes, fused and stacks routes), the mixed path, the
batch graph, plugin-level filters on device frames and on host frames, and
the VapourSynth plugin entry (setFilterError through the mock host).
"""
import numpy as np
import pytest

from bm3d_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

MESSAGE = {1: "out of memory", 2: "out of memory", 3: "launch failure"}
SITES = {"device_alloc": 1, "host_alloc": 2, "launch": 3}
PRM = dict(block_step=4, bm_range=4, radius=1, ps_num=2, ps_range=2, extractor_exp=10)


def _lib():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import _capi as C
    return C


def sweep(C, site, run, expect, limit=400):
    """Fails event k of `run` for k = 0, 1, ... (a fresh `run` each time,
    then once more undisturbed on the same objects); returns the number of
    failing calls seen before a call completed with the countdown pending."""
    lib = C.lib()
    failures = 0
    try:
        for k in range(limit):
            C.check(lib.bm3d_debug_inject_failure(site, k))
            try:
                got = run()
            except C.BM3DError as e:
                assert e.code == C.BM3D_EHIP, (k, e.code, str(e))
                assert MESSAGE[site] in str(e), (k, str(e))
                failures += 1
                expect(run())  # the injection is one-shot: the retry is undisturbed
                continue
            expect(got)
            return failures
        raise AssertionError(f"no call completed within {limit} events")
    finally:
        lib.bm3d_debug_inject_failure(site, -1)


def _clip(N=5, H=40, W=56, seed=500):
    return synth.make_clip(H, W, N, 1, noise_sd=10.0, seed=seed, motion=True)


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
@pytest.mark.parametrize("site", ["device_alloc", "launch"])
def test_plans(site, precision):
    C = _lib()
    from bm3d_amd import core
    _, noisy = _clip()
    N, _, H, W = noisy.shape
    src = torch.from_numpy(noisy).cuda()
    kw = dict(PRM, precision=precision)
    pb0, pf0 = core.Plan(W, H, 10.0, **kw), core.Plan(W, H, 10.0, final=True, **kw)
    want_b = pb0.execute_vagg(src)
    want_f = pf0.execute_vagg(src, want_b)
    want_s = pf0.execute(src, want_b)

    def run():
        pb = core.Plan(W, H, 10.0, **kw)
        pf = core.Plan(W, H, 10.0, final=True, **kw)
        state["plans"] = (pb, pf)
        b = pb.execute_vagg(src)
        return b, pf.execute_vagg(src, b), pf.execute(src, b)

    def expect(got):
        torch.cuda.synchronize()
        assert torch.equal(got[0], want_b) and torch.equal(got[1], want_f) and torch.equal(got[2], want_s)

    state = {}
    n = sweep(C, SITES[site], run, expect)
    print(f"{site} {precision}: {n} failing calls")
    assert n >= 3, n  # plan set-up and buffers, or the launches of 3 calls
    # the plans of the last failed call recover too
    if state:
        pb, pf = state["plans"]
        b = pb.execute_vagg(src)
        assert torch.equal(b, want_b) and torch.equal(pf.execute_vagg(src, b), want_f)


@pytest.mark.parametrize("site", ["device_alloc", "launch"])
def test_batch(site):
    C = _lib()
    from bm3d_amd import core
    clips = [_clip(N=4, seed=510 + k)[1] for k in range(3)]
    srcs = [torch.from_numpy(c).cuda() for c in clips]

    def jobs():
        return [dict(PRM, src=s, passes=2, sigma=[8.0 + 2 * k]) for k, s in enumerate(srcs)]

    ref = core.Batch(jobs())
    ref.execute()
    torch.cuda.synchronize()
    want = [j["out"].clone() for j in ref.jobs]
    del ref

    def run():
        b = core.Batch(jobs())
        b.execute()
        torch.cuda.synchronize()
        return b

    def expect(b):
        for j, w in zip(b.jobs, want):
            assert torch.equal(j["out"], w)

    n = sweep(C, SITES[site], run, expect)
    print(f"{site}: {n} failing calls")
    assert n >= 2, n


@pytest.mark.parametrize("site", ["device_alloc", "host_alloc", "launch"])
def test_filters_device_and_host_frames(site):
    C = _lib()
    from bm3d_amd import vs
    _, noisy = _clip(N=8)
    kw = dict(PRM, sigma=[10.0])
    dev = vs.Clip.from_array(noisy, "GRAY")
    want = vs.BM3Dv2(dev, ref=vs.BM3Dv2(dev, **kw), **kw).numpy()
    host = vs.Clip.from_array(noisy, "GRAY", device="cpu")
    routes = [host] if site == "host_alloc" else [dev, host]
    for clip in routes:
        def run():
            return vs.BM3Dv2(clip, ref=vs.BM3Dv2(clip, **kw), **kw).numpy()

        def expect(got):
            np.testing.assert_array_equal(got, want)

        n = sweep(C, SITES[site], run, expect)
        print(f"{site} {'host' if clip.on_host else 'device'}: {n} failing calls")
        assert n >= 1, n
