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
    def __init__(self, slots, t, fp=1, cost=None):
        self.footprint = fp
        self.t_arrival = t
        self.cost = slots if cost is None else cost

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


def test_select_batch_by_cost():
    """With a cost cap the batch is sized by summed cost: small requests fill it up to
    the cap, a request costlier than the cap runs alone, and the count cap still holds."""
    from gbasis_amd.serving import select_batch

    small = [_Req(10, 0.0, cost=100.0) for _ in range(10)]
    big = _Req(10, 0.0, cost=5000.0)
    pend = small + [big]
    take = select_batch(pend, 0.0, 1e9, 16, "sjf", batch_cost=450.0)
    assert take == small[:4]
    assert select_batch(pend, 0.0, 1e9, 3, "sjf", batch_cost=450.0) == small[:3]
    # the big request, once overdue (and first to arrive), opens a batch of its own
    late = [big] + [_Req(10, 0.9, cost=100.0) for _ in range(3)]
    assert select_batch(late, 1.0, 1e9, 16, "sjf", max_wait=0.5, batch_cost=450.0) == [big]
    # a cap above the big request's cost lets small ones join it
    take = select_batch(late, 1.0, 1e9, 16, "sjf", max_wait=0.5, batch_cost=5200.0)
    assert take[0] is big and len(take) == 3        # 5000 + 2 * 100 <= 5200


def test_estimated_flops_equals_quartet_sum():
    """The grouped FLOP estimate equals the sum of the FLOP model over every unique
    shell quartet of a molecule, enumerated directly."""
    from gbasis_amd.roofline import class_flops
    from gbasis_amd.serving import estimated_flops
    from gbasis_amd.tables import MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    m = MoleculeTables(batch_molecule(1003, n_atoms=3))
    # the FLOP model counts a class in the reference's order, which leaves bra and ket
    # of equal l_a unordered: the pair with the smaller (l_max, l_min) goes first, as
    # in the estimate
    key = [(max(a, b), min(a, b)) for a, b in zip(m.pair_la.tolist(), m.pair_lb.tolist())]
    total = 0.0
    for p in range(m.npair):
        for q in range(p + 1):
            x, y = sorted((key[p], key[q]))
            fp, fc = class_flops(*x, *y)
            total += float(m.pair_nprim[p]) * float(m.pair_nprim[q]) * fp + fc
    assert abs(estimated_flops(m) - total) <= 1e-9 * total
