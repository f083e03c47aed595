"""CPU checks of the mixed matcher's key selection (csrc/bm3d_match_dx.hip).

The matcher keeps the K smallest 32-bit keys of a window by merging 8 new
keys at a time: the 8 are sorted with a comparator network (sort8), paired
as min(k[i], b[K-1-i]) with the kept list, and the result is sorted by a
bitonic merge padded at the front with zeros (KeyList::push8).

* sort8's comparators are read from the kernel source and checked by the
  0-1 principle: a comparator network sorts every input iff it sorts all
  2^8 inputs of zeros and ones.
* The merge is restated step by step and checked against sorted(keys)[:K]
  for every K the kernel allows (4 and 8..16), with random and heavily tied
  keys, over whole windows of 8-key batches.
"""
import os
import random
import re

import pytest

SRC = os.path.join(os.path.dirname(__file__), os.pardir, "vapoursynth-bm3dcuda_amd", "csrc", "bm3d_match_dx.hip")


def _source():
    with open(SRC) as f:
        return f.read()


def _sort8_network():
    src = _source()
    body = src[src.index("void sort8("):]
    body = body[:body.index("\n}\n")]
    pairs = [(int(a), int(b)) for a, b in re.findall(r"cex\(v\[(\d)\],\s*v\[(\d)\]\)", body)]
    assert pairs, "sort8 not found"
    return pairs


def _apply(net, v):
    v = list(v)
    for a, b in net:
        if v[a] > v[b]:
            v[a], v[b] = v[b], v[a]
    return v


def test_sort8_network_sorts_every_input():
    net = _sort8_network()
    assert len(net) == 19
    assert all(a < b for a, b in net)
    for m in range(256):
        bits = [(m >> i) & 1 for i in range(8)]
        assert _apply(net, bits) == sorted(bits), bits


def _cex_all(v, pairs):
    for a, b in pairs:
        if v[a] > v[b]:
            v[a], v[b] = v[b], v[a]
    return v


def _push8_k4(k, b):
    """KeyList<4>::push8: the 4 smallest of b by sorted halves and the
    min(lo[i], hi[3-i]) pairing, then the same pairing into k."""
    sort4 = [(0, 1), (2, 3), (0, 2), (1, 3), (1, 2)]
    bitonic4 = [(0, 2), (1, 3), (0, 1), (2, 3)]
    b = _cex_all(list(b[:4]), sort4) + _cex_all(list(b[4:]), sort4)
    c = _cex_all([min(b[i], b[7 - i]) for i in range(4)], bitonic4)
    return _cex_all([min(k[i], c[3 - i]) for i in range(4)], bitonic4)


def _push8(k, b, net):
    """KeyList<K>::push8 for K >= 8: returns the new ascending list."""
    K = len(k)
    if K == 4:
        return _push8_k4(k, b)
    P = 8 if K <= 8 else 16
    b = _apply(net, b)
    c = [0] * (P - K) + [min(k[i], b[K - 1 - i]) if K - 1 - i < 8 else k[i] for i in range(K)]
    d = P // 2
    while d >= 1:
        for i in range(P):
            if (i & d) == 0 and c[i] > c[i + d]:
                c[i], c[i + d] = c[i + d], c[i]
        d //= 2
    return c[P - K:]


@pytest.mark.parametrize("K", [4] + list(range(8, 17)))
@pytest.mark.parametrize("spread", [1 << 32, 40, 3])
def test_push8_keeps_the_k_smallest(K, spread):
    net = _sort8_network()
    rng = random.Random(K * 1000 + spread % 997)
    for _ in range(60):
        k = [0xFFFFFFFF] * K  # KeyList::init
        seen = []
        for _ in range(rng.randint(1, 6)):
            batch = [rng.randrange(spread) for _ in range(8)]
            # out-of-plane candidates arrive as all-ones keys
            for i in range(8):
                if rng.random() < 0.1:
                    batch[i] = 0xFFFFFFFF
            seen += batch
            k = _push8(k, batch, net)
            assert k == sorted(seen + [0xFFFFFFFF] * K)[:K]


def test_kernel_key_count_is_mergeable():
    m = re.search(r"constexpr int kKSpatial = (\d+);", _source())
    assert m and 8 <= int(m.group(1)) <= 16


def _insert(k, x):
    """KeyList::insert: k[0] = min(k[0], x), k[i] = the median of
    (k[i-1], x, k[i]) for the old k (the clamp of x to [k[i-1], k[i]])."""
    out = [min(k[0], x)]
    for i in range(1, len(k)):
        out.append(sorted((k[i - 1], x, k[i]))[1])
    return out


def test_insert_reads_median_of_three_in_source():
    """The kernel's insertion computes max(min(prev, key), min(max(prev,
    key), cur)), the median of three (the form the compiler emits as one
    v_med3_u32), from the list's old entries."""
    src = _source()
    body = src[src.index("void insert(uint32_t key)"):]
    body = body[:body.index("\n    }\n")]
    assert "k[0] = min(k[0], key)" in body
    assert "max(min(prev, key), min(max(prev, key), cur))" in body
    assert re.search(r"const uint32_t cur = k\[i\];", body) and "prev = cur;" in body


@pytest.mark.parametrize("K", [4, 8, 10, 16])
def test_insert_keeps_the_k_smallest(K):
    """One key at a time, the list is sorted(keys so far)[:K] (all ones
    marking empty entries), for random, tied and descending keys."""
    rng = random.Random(K)
    empty = 0xFFFFFFFF
    for trial in range(200):
        n = rng.randrange(1, 60)
        if trial % 3 == 0:
            keys = [rng.randrange(8) for _ in range(n)]
        elif trial % 3 == 1:
            keys = sorted((rng.randrange(1 << 32) for _ in range(n)), reverse=True)
        else:
            keys = [rng.randrange(1 << 32) for _ in range(n)]
        k = [empty] * K
        for j, x in enumerate(keys):
            k = _insert(k, x)
            want = (sorted(keys[:j + 1]) + [empty] * K)[:K]
            assert k == want
