"""A caller may capture a plan's bm3d_execute_vagg into a HIP graph (torch's
CUDA graph API on ROCm), after one eager call has sized the plan's buffers.
The call forks its input-range and zero-fill kernels onto a second stream
and joins it back before filtering; the capture must hold that fork and
join. Bar: every replay bit-identical to the eager result (extractor_exp
set), with eager work and synchronisations between the replays."""
import pytest

from bm3d_amd import synth

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _core():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from bm3d_amd import core
    return core


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_execute_vagg_graph_replays_equal_eager(precision):
    core = _core()
    H, W, N = 48, 64, 6
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=10.0, seed=96, motion=True)
    src = torch.from_numpy(noisy).cuda()
    args = (W, H, 10.0, 4, 4, 1, 2, 2)

    def plans():
        return (core.Plan(*args, precision=precision, extractor_exp=10),
                core.Plan(*args, final=True, precision=precision, extractor_exp=10))

    eb, ef = plans()
    want_b = eb.execute_vagg(src)
    want_f = ef.execute_vagg(src, want_b)
    torch.cuda.synchronize()

    pb, pf = plans()
    b, f = torch.empty_like(want_b), torch.empty_like(want_f)
    pb.execute_vagg(src, out=b)  # eager first: buffers and the second stream
    pf.execute_vagg(src, b, out=f)
    torch.cuda.synchronize()
    assert torch.equal(b, want_b) and torch.equal(f, want_f)

    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        pb.execute_vagg(src, out=b)
        pf.execute_vagg(src, b, out=f)
    for rep in range(3):
        b.fill_(float("nan"))
        f.fill_(float("nan"))
        other = eb.execute_vagg(src)  # eager work of the program between replays
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(b, want_b), rep
        assert torch.equal(f, want_f), rep
        assert torch.equal(other, want_b), rep


def test_captured_plan_is_pinned():
    """A call made while the stream captures pins the plan: the graph holds
    its buffers and tiles, so later calls that would grow the buffers, and
    tuning changes, fail with an error instead of moving them. Calls that
    fit the buffers keep working, eager and replayed."""
    core = _core()
    from bm3d_amd import _capi as C
    H, W, N = 40, 48, 6
    _, noisy = synth.make_clip(H, W, N, 1, noise_sd=10.0, seed=97, motion=True)
    src = torch.from_numpy(noisy).cuda()
    p = core.Plan(W, H, 10.0, 4, 4, 1, 2, 2, extractor_exp=10)
    want = p.execute_vagg(src, first=0, count=2)
    torch.cuda.synchronize()
    out = torch.empty_like(want)
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        p.execute_vagg(src, first=0, count=2, out=out)
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(out, want)
    with pytest.raises(C.BM3DError, match="captured graph"):
        p.set_tuning(filter_tile_x=4)
    with pytest.raises(C.BM3DError, match="captured graph"):
        p.autotune(src, first=0, count=2)
    with pytest.raises(C.BM3DError, match="captured graph"):
        p.execute_vagg(src)  # all 6 frames: larger buffers than the capture saw
    again = p.execute_vagg(src, first=0, count=2)  # fits: still fine
    g.replay()
    torch.cuda.synchronize()
    assert torch.equal(again, want) and torch.equal(out, want)
