"""bm3d_execute_noalloc: bm3d_execute without device allocation, its buffers
in a static 64 MiB device array. Called through the C ABI; bar: stacks
bit-identical to bm3d_execute's on the same plan and inputs (extractor_exp
set), on two calls in a row -- the second with a smaller input -- for the FP32
and mixed-precision kernels; a plan that needs more than the array for one
evaluation is refused with BM3D_EINVAL naming the size."""
import ctypes

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


def _noalloc(C, plan, src, ref=None):
    n = src.shape[0]
    out = plan.new_stacks(n, src.device)
    out.fill_(float("nan"))
    from bm3d_amd import core
    s = core.clip_desc(src)
    r = core.clip_desc(ref) if ref is not None else None
    C.check(C.lib().bm3d_execute_noalloc(plan._h, ctypes.c_void_p(out.data_ptr()), 0, 0, ctypes.byref(s),
                                         ctypes.byref(r) if r else None, 0, n, None))
    return out


@pytest.mark.parametrize("precision", ["fp32", "mixed"])
def test_noalloc_equals_execute_twice(precision):
    core = _core()
    from bm3d_amd import _capi as C
    cases = [(96, 128, 6, dict(sigma=12.0, block_step=4, bm_range=6, radius=1, ps_num=2, ps_range=3)),
             (40, 56, 3, dict(sigma=20.0, block_step=3, bm_range=4, radius=2, ps_num=2, ps_range=2))]
    for k, (H, W, N, prm) in enumerate(cases):  # the second call's input is smaller
        _, noisy = synth.make_clip(H, W, N, 1, noise_sd=prm["sigma"], seed=60 + k, motion=True)
        src = torch.from_numpy(noisy).cuda()
        for final in (False, True):
            plan = core.Plan(W, H, final=final, extractor_exp=10, precision=precision, **prm)
            ref = src.flip(0).contiguous() if final else None
            want = plan.execute(src, ref)
            got = _noalloc(C, plan, src, ref)
            torch.cuda.synchronize()
            assert torch.equal(got, want), (k, final)


def test_noalloc_refuses_large_plans():
    core = _core()
    from bm3d_amd import _capi as C
    plan = core.Plan(3840, 2160, 20.0, block_step=4, bm_range=16, radius=2, ps_num=2, ps_range=6)
    src = torch.zeros((1, 1, 2160, 3840), device="cuda")
    with pytest.raises(C.BM3DError, match="bytes of scratch"):
        _noalloc(C, plan, src)
