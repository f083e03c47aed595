"""bm3d_execute_window: the reference's packed window layout through the C ABI.

The reference plugin stages one window per call in
d_src = [(final ? 2 : 1)][planes][2r+1][height][stride], the basic
estimate's slots first in the final pass (source/kernel.cu:43-44,
source/source.cpp:244-262), and its graph writes one stack
d_res = [planes][2r+1][2][height][stride] (source/source.cpp:362-366).
bm3d_execute_window is the call a maintainer swaps in for that graph launch.
Bar: stacks bit-identical to the oracle's window evaluation (extractor_exp
set, so every sum is exact), with a row pitch wider than the plane, for the
basic and final passes, one plane and three (chroma), radius 0, 1 and 2.
"""
import ctypes

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


CASES = [
    # (radius, chroma, sigma)
    (0, False, [12.0]),
    (1, False, [15.0]),
    (1, True, [10.0, 8.0, 12.0]),
    (2, False, [20.0]),
]


@pytest.mark.parametrize("radius,chroma,sigma", CASES, ids=[f"r{r}-{'yuv' if c else 'gray'}" for r, c, _ in CASES])
@pytest.mark.parametrize("final", [False, True], ids=["basic", "final"])
def test_execute_window_matches_oracle(radius, chroma, sigma, final):
    core = _core()
    from bm3d_amd import _capi as C
    H, W, N, stride = 40, 52, 5, 57
    P = 3 if chroma else 1
    E = 10
    _, noisy = synth.make_clip(H, W, N, P, noise_sd=sigma[0], seed=31 + radius, motion=True, yuv=chroma)
    basic = np.clip(noisy + np.float32(0.01), 0, 1).astype(np.float32)  # any plausible ref will do
    n = 2
    tw = 2 * radius + 1
    slots = [min(max(n - radius + z, 0), N - 1) for z in range(tw)]
    win = np.stack([noisy[f] for f in slots], axis=1)  # [planes][2r+1][H][W]
    bwin = np.stack([basic[f] for f in slots], axis=1)
    packed = np.zeros(((2 if final else 1), P, tw, H, stride), np.float32)
    if final:
        packed[0, :, :, :, :W] = bwin
        packed[1, :, :, :, :W] = win
    else:
        packed[0, :, :, :, :W] = win
    plan = core.Plan(W, H, sigma, 4, 4, radius, 2, 2, chroma=chroma, final=final, extractor_exp=E, stride=stride)
    d_src = torch.from_numpy(packed).cuda()
    d_res = torch.full((P, tw, 2, H, stride), float("nan"), device="cuda")
    rc = C.lib().bm3d_execute_window(None, ctypes.c_void_p(d_res.data_ptr()),
                                     ctypes.c_void_p(d_src.data_ptr()), None)
    assert rc != 0 and C.lib().bm3d_last_error()
    C.check(C.lib().bm3d_execute_window(plan._h, ctypes.c_void_p(d_res.data_ptr()),
                                        ctypes.c_void_p(d_src.data_ptr()), None))
    assert C.lib().bm3d_last_error() in (b"", "")  # cleared by the successful call
    torch.cuda.synchronize()
    got = d_res.cpu().numpy()[..., :W]
    sig = [O.scale_sigma(s, final) for s in (sigma + [sigma[-1]] * 3)[:P]]
    want = O.bm3d_window(win, bwin if final else None, sig, 4, 4, radius, 2, 2, chroma, final,
                         float(2 ** E))
    np.testing.assert_array_equal(got, want)
