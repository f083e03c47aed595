"""Thread-parallel driver of the CPU oracle for the larger GPU test cases:
the per-frame window evaluations of one pass run concurrently (the oracle's
C entry points release the interpreter lock), then VAggregate and the second
pass exactly as oracle.bm3dv2 / oracle.two_pass do."""
import concurrent.futures as cf
import os

import numpy as np

from oracle import oracle as O

THREADS = min(16, os.cpu_count() or 1)


def bm3dv2(clip, ref, p: O.Params, dtype=np.float32):
    N, npl = clip.shape[:2]
    sig = [float(s) for s in (list(p.sigma) + [p.sigma[-1]] * 3)[:3]]
    process = [s >= O.FLT_EPSILON for s in sig][:npl]
    if not any(process):
        return np.array(clip, dtype=dtype, copy=True)
    with cf.ThreadPoolExecutor(THREADS) as ex:
        stacks = list(ex.map(lambda n: O.bm3d_frame(clip, ref, n, p, dtype), range(N)))
    if p.radius == 0:
        return np.stack(stacks)
    r = p.radius
    return np.stack([O.vaggregate([stacks[min(max(n - r + i, 0), N - 1)] for i in range(2 * r + 1)],
                                  clip[n], n, N, r, process, dtype) for n in range(N)])


def two_pass(clip, p: O.Params, dtype=np.float32):
    basic = bm3dv2(clip, None, p, dtype)
    return basic, bm3dv2(clip, basic.astype(dtype), p, dtype)
