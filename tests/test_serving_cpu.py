"""Host logic of the ERI server that needs no GPU: the output arena allocator."""
import numpy as np

from gbasis_amd.serving import Arena


def test_arena_first_fit_alignment_and_coalescing():
    ar = Arena(1024)
    a = ar.alloc(100)
    b = ar.alloc(64)
    c = ar.alloc(1)
    assert a == (0, 128) and b == (128, 64) and c == (192, 64)
    assert ar.alloc(2000) is None
    ar.release(*b)
    assert ar.alloc(64) == (128, 64)        # first fit reuses the hole
    ar.release(128, 64)
    ar.release(*a)
    ar.release(*c)
    assert ar.free == [(0, 1024)]           # everything coalesced back


def test_arena_random_alloc_release_never_overlaps():
    rng = np.random.default_rng(0)
    ar = Arena(1 << 16)
    live = {}
    for step in range(2000):
        if live and rng.random() < 0.45:
            k = list(live)[rng.integers(len(live))]
            ar.release(*live.pop(k))
        else:
            r = ar.alloc(int(rng.integers(1, 4000)))
            if r is not None:
                live[step] = r
        spans = sorted(live.values()) + sorted(ar.free)
        spans.sort()
        pos = 0
        for off, n in spans:                 # live and free regions tile the arena
            assert off == pos and n > 0
            pos += n
        assert pos == ar.numel


class _Req:
    def __init__(self, slots, t, fp=1):
        self.footprint = fp
        self.t_arrival = t

        class _T:
            n_slots = slots
        self.tables = _T()


def test_select_batch_fifo_and_sjf_with_aging():
    from gbasis_amd.serving import select_batch

    big, small1, small2, mid = _Req(10000, 0.9), _Req(10, 0.9), _Req(12, 0.95), _Req(15, 0.97)
    pend = [big, small1, small2, mid]
    assert select_batch(pend, 1.0, 100, 8, "fifo") == pend
    assert select_batch(pend, 1.0, 2, 8, "fifo") == [big, small1]
    # smallest cost first; the batch keeps to requests within 2x of its opener
    assert select_batch(pend, 1.0, 100, 8, "sjf", age_tau=0.1) == [small1, small2, mid]
    # a request waiting longer than max_wait opens the batch
    assert select_batch(pend, 2.0, 100, 8, "sjf", age_tau=0.1, max_wait=0.5)[0] is big
    # memory: the best-scored request waits rather than being overtaken
    assert select_batch([_Req(10, 0.0, fp=5), _Req(20, 0.0)], 0.0, 4, 8, "sjf") == []
