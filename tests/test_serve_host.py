"""Host-side serving logic (no device): the arena free list."""
import numpy as np

from cuvarbase_amd.serve import FreeList


def test_free_list_alloc_release_coalesce():
    fl = FreeList(10 * 256)
    a = fl.alloc(300)          # rounds to 512
    b = fl.alloc(256)
    c = fl.alloc(1000)         # 1024
    assert a == (0, 512) and b == (512, 256) and c == (768, 1024)
    assert fl.used == 1792
    assert fl.alloc(4096) is None
    fl.release(b)
    assert fl.alloc(200) == (512, 256)   # first fit reuses the hole
    fl.release((512, 256))
    fl.release(a)
    fl.release(c)
    assert fl.used == 0
    assert fl._free == [(0, 10 * 256)]
    assert fl.largest_free == 2560


def test_free_list_random_churn_never_overlaps():
    rng = np.random.default_rng(0)
    fl = FreeList(1 << 20)
    live = []
    for _ in range(2000):
        if live and rng.random() < 0.45:
            fl.release(live.pop(rng.integers(len(live))))
        else:
            blk = fl.alloc(int(rng.integers(1, 40000)))
            if blk is not None:
                live.append(blk)
        spans = sorted(live)
        for (o1, s1), (o2, _) in zip(spans, spans[1:]):
            assert o1 + s1 <= o2
        assert fl.used == sum(s for _, s in live)
    for blk in live:
        fl.release(blk)
    assert fl._free == [(0, 1 << 20)]
