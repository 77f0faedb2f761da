"""Continuous-batching ERI server on one GPU.

A request is one molecule (a list of shells) and a precision ("fp64" or
"mixed"); its answer is the molecule's canonical 8-fold packed ERI array,
left on the device (a view that the caller may copy to the host; a caller
that uses it on another stream records that stream on it first, as for any
torch tensor shared between streams).

* Host tables are built when a request is submitted.
* Admission is by HBM footprint: packed output, work queue and tables of every
  request on the device are charged against a byte budget; a request is taken
  from the FIFO queue only when it fits next to the requests in flight (and its
  batch's packed output fits in the output arena), and its bytes are returned by
  ``release``.
* Memory: packed outputs are carved first-fit from one arena allocated when the
  server starts, and the engine's work-queue and accumulator buffers are
  grow-only per stream, so steady-state serving allocates and frees no device
  memory (a device free would synchronise the GPU and serialise the streams).
  A result view is valid until its request is released.
* Continuous batching: every stream that has no batch in flight takes
  admitted requests that fit as one batch -- one table upload, one Schwarz pass,
  one device-built work queue whose launches read their counts on the device,
  every class launch -- so batches launched while others run overlap on the
  device.  FP64 and mixed requests share a batch through per-molecule FP32
  cutoffs.
* Order: with policy "sjf" (default) the waiting request with the smallest
  cost / (1 + age / age_tau) opens a batch (cost: the algorithmic FLOPs of all of
  the molecule's shell quartets, ``estimated_flops``; age: time since arrival), and
  the batch takes the next requests by the same score while its summed cost stays
  within max(opener's cost, batch_flops) -- batches are sized by work, not by
  count, so that a batch of small molecules and one large molecule take similar
  times, and a request waiting longer than max_wait (0.5 s) opens the next batch.
  Policy "fifo" takes requests in arrival order.
* Results do not depend on the company a request keeps: the engine runs in
  deterministic mode (variant choices depend on class and primitive bin only)
  and every packed slot is produced by one shell quartet whose arithmetic reads
  only its own molecule's tables, so a request's packed array is bit-identical
  alone and in any batch.

Calls on one server are not thread-safe; drive it from one host thread.
"""

import collections
import time

import numpy as np
import torch

from .engine import DeviceTables, EriEngine
from .tables import BatchTables, MoleculeTables

PRECISIONS = ("fp64", "mixed")


class EriRequest:
    """One molecule's ERI request and, once done, its packed result (device)."""

    def __init__(self, rid, basis, precision, t_arrival):
        if precision not in PRECISIONS:
            raise ValueError(f"precision must be one of {PRECISIONS}")
        self.id = rid
        self.precision = precision
        self.t_arrival = t_arrival
        self.tables = MoleculeTables(basis)
        m = self.tables
        n_cand = m.npair * (m.npair + 1) // 2
        # packed output + queue room (up to four sub-quartets per candidate, both
        # precision variants) + pair records (first- and second-centre copies)
        self.footprint = int(8 * m.n_slots + 64 * n_cand
                             + 80 * (m.prim_pairs.shape[0] + m.sub_prims.shape[0])
                             + 3 * 64 * m.npair + (1 << 20))
        self.cost = estimated_flops(m)
        self.t_launch = None
        self.t_done = None
        self.result = None
        self.released = False
        self._batch = None

    @property
    def nfunc(self):
        return self.tables.nfunc

    @property
    def done(self):
        return self.t_done is not None

    def latency(self):
        return None if self.t_done is None else self.t_done - self.t_arrival


def estimated_flops(m):
    """Algorithmic FLOPs (gbasis_amd.roofline) of every unique shell quartet of the
    molecule tables ``m``, before screening: pairs are grouped by their canonical
    (l_a, l_b); a group pair (P, Q) holds n_P n_Q quartets (n_P (n_P + 1) / 2 when
    P = Q) with sum_P K sum_Q K primitive quartets ((sum_P K)^2 + sum_P K^2) / 2
    when P = Q), K the primitive pairs of a pair."""
    from .roofline import class_flops

    la = np.maximum(m.pair_la, m.pair_lb)
    lb = np.minimum(m.pair_la, m.pair_lb)
    keys, inv = np.unique(la * 8 + lb, return_inverse=True)
    k = m.pair_nprim.astype(np.float64)
    n = np.bincount(inv).astype(np.float64)
    s1 = np.bincount(inv, weights=k)
    s2 = np.bincount(inv, weights=k * k)
    total = 0.0
    for i in range(keys.size):
        for j in range(i, keys.size):
            fp, fc = class_flops(keys[i] // 8, keys[i] % 8, keys[j] // 8, keys[j] % 8)
            if i == j:
                prim, quart = 0.5 * (s1[i] * s1[i] + s2[i]), 0.5 * n[i] * (n[i] + 1)
            else:
                prim, quart = s1[i] * s1[j], n[i] * n[j]
            total += prim * fp + quart * fc
    return total


class _Batch:
    def __init__(self, reqs, stream, dev, out, event, plans, arena_off, arena_len):
        self.reqs = reqs
        self.stream = stream
        self.dev = dev
        self.out = out
        self.event = event
        self.plans = plans
        self.arena = (arena_off, arena_len)
        self.live = len(reqs)


class Arena:
    """First-fit sub-allocation of one device buffer (offsets in elements)."""

    ALIGN = 64   # elements (512 bytes)

    def __init__(self, numel):
        self.numel = int(numel)
        self.free = [(0, self.numel)]   # sorted, coalesced (offset, length)

    def alloc(self, n):
        n = -(-max(int(n), 1) // self.ALIGN) * self.ALIGN
        for i, (o, ln) in enumerate(self.free):
            if ln >= n:
                if ln == n:
                    self.free.pop(i)
                else:
                    self.free[i] = (o + n, ln - n)
                return o, n
        return None

    def release(self, off, n):
        self.free.append((off, n))
        self.free.sort()
        merged = []
        for o, ln in self.free:
            if merged and merged[-1][0] + merged[-1][1] == o:
                merged[-1] = (merged[-1][0], merged[-1][1] + ln)
            else:
                merged.append((o, ln))
        self.free = merged


class EriServer:
    """Admission, continuous batching and completion of ERI requests."""

    def __init__(self, device="cuda", threshold=1e-12, fp32_cutoff=1e-2, hbm_budget=None,
                 n_streams=4, max_batch=32, clock=time.perf_counter, policy="sjf",
                 age_tau=0.1, max_wait=0.5, batch_flops=8e12):
        if policy not in ("sjf", "fifo"):
            raise ValueError("policy must be 'sjf' or 'fifo'")
        self.policy = policy
        self.age_tau = float(age_tau)
        self.max_wait = float(max_wait)
        self.batch_flops = None if batch_flops is None else float(batch_flops)
        # an engine of its own: deterministic variant choices and per-stream workspaces
        self.eng = EriEngine(device)
        self.eng.deterministic = True
        self.eng.stream_workspaces = True
        self.device = torch.device(device)
        self.threshold = float(threshold)
        self.fp32_cutoff = float(fp32_cutoff)
        if hbm_budget is None:
            free, _ = torch.cuda.mem_get_info(self.device)
            hbm_budget = int(0.4 * free)
        self.budget = int(hbm_budget)
        self.used = 0
        # packed outputs of every admitted request fit: each footprint counts them
        self.arena = Arena(self.budget // 8)
        self.arena_buf = torch.empty(self.arena.numel, dtype=torch.float64, device=self.device)
        self.max_batch = int(max_batch)
        # streams at the default priority
        self.streams = [torch.cuda.Stream(self.device) for _ in range(int(n_streams))]
        self.inflight = [None] * len(self.streams)
        self.pending = collections.deque()
        self.clock = clock
        self._next_id = 0
        self.batches_launched = 0
        # host seconds spent building request tables (submit) and enqueueing batches
        self.host_seconds = {"tables": 0.0, "launch": 0.0}

    # ------------------------------------------------------------------
    def submit(self, basis, precision="fp64", t_arrival=None, dispatch=True):
        """Queue a request (host tables are built here) and, with ``dispatch``, launch
        what fits on idle streams.  Completions are reported by ``step`` only."""
        t = self.clock() if t_arrival is None else t_arrival
        h0 = time.perf_counter()
        req = EriRequest(self._next_id, basis, precision, t)
        self.host_seconds["tables"] += time.perf_counter() - h0
        self._next_id += 1
        if req.footprint > self.budget:
            raise ValueError(f"request needs {req.footprint} bytes of HBM, more than the "
                             f"server's budget of {self.budget}")
        self.pending.append(req)
        if dispatch:
            self._dispatch()
        return req

    def step(self):
        """Complete finished batches, then launch new ones on idle streams.
        Returns the requests completed by this step."""
        done = self._poll()
        self._dispatch()
        return done

    def release(self, req):
        """Return a completed request's HBM to the budget (drops the server's view)."""
        if req.released:
            return
        if not req.done:
            raise RuntimeError("request is not complete")
        req.released = True
        req.result = None
        self.used -= req.footprint
        b = req._batch
        req._batch = None
        b.live -= 1
        if b.live == 0:
            self.arena.release(*b.arena)

    def busy(self):
        return bool(self.pending) or any(b is not None for b in self.inflight)

    def drain(self, on_done=None, spin_s=50e-6):
        """Step until every submitted request is complete."""
        while self.busy():
            done = self.step()
            if on_done is not None:
                for r in done:
                    on_done(r)
            if not done:
                time.sleep(spin_s)

    # ------------------------------------------------------------------
    def _poll(self):
        done = []
        for i, b in enumerate(self.inflight):
            if b is not None and b.event.query():
                t = self.clock()
                for r, off in zip(b.reqs, b.dev.batch.mol_slot_off):
                    r.result = b.out[off:off + r.tables.n_slots]
                    r.t_done = t
                    r._batch = b
                    done.append(r)
                b.dev = b.plans = None
                self.inflight[i] = None
        return done

    def _dispatch(self):
        for i, b in enumerate(self.inflight):
            if b is not None or not self.pending:
                continue
            take = select_batch(list(self.pending), self.clock(), self.budget - self.used,
                                self.max_batch, self.policy, self.age_tau, self.max_wait,
                                self.batch_flops)
            for r in take:
                self.pending.remove(r)
            need = sum(r.footprint for r in take)
            slots = sum(r.tables.n_slots for r in take)
            while take:
                region = self.arena.alloc(slots)
                if region is not None:
                    break
                # fragmented arena: hand the last request back and try a smaller batch
                r = take.pop()
                self.pending.appendleft(r)
                need -= r.footprint
                slots -= r.tables.n_slots
            if not take:
                return   # head of the queue waits for memory to be released
            self.used += need
            h0 = time.perf_counter()
            self.inflight[i] = self._launch(take, self.streams[i], region)
            self.host_seconds["launch"] += time.perf_counter() - h0

    def _launch(self, reqs, stream, region):
        eng = self.eng
        t = self.clock()
        for r in reqs:
            r.t_launch = t
        with torch.cuda.stream(stream):
            batch = BatchTables([r.tables for r in reqs])
            mixed = any(r.precision == "mixed" for r in reqs)
            cut = (np.array([self.fp32_cutoff if r.precision == "mixed" else 0.0 for r in reqs])
                   if mixed else None)
            dev = DeviceTables(batch, self.device, mol_fp32_cutoff=cut)
            need_bound = self.threshold > 0 or mixed
            bound = eng.pair_bounds(dev, eng.schwarz(dev)) if need_bound else None
            plans = eng.plan_device(dev, bound, self.threshold, 0.0, stats=False,
                                    device_counts=True)
            out = self.arena_buf[region[0]:region[0] + batch.n_slots]
            out.zero_()
            eng.run(dev, plans, out)
            ev = torch.cuda.Event()
            ev.record(stream)
        self.batches_launched += 1
        return _Batch(reqs, stream, dev, out, ev, plans, *region)


def select_batch(pending, now, budget_left, max_batch, policy="sjf", age_tau=0.1,
                 max_wait=0.5, batch_cost=None):
    """The requests of the next batch, from ``pending`` (arrival order).  Objects need
    ``footprint``, ``t_arrival`` and ``tables.n_slots``, and ``cost`` when
    ``batch_cost`` is given.  With "sjf" a request that has waited longer than
    ``max_wait`` seconds opens the batch (oldest first).  The batch then holds at most
    ``max_batch`` requests and, with ``batch_cost``, a summed cost of at most
    max(opener's cost, batch_cost); without it, requests of at most twice the
    opener's packed slots."""
    take, need = [], 0
    if policy == "fifo":
        for r in pending:
            if len(take) >= max_batch or need + r.footprint > budget_left:
                break
            take.append(r)
            need += r.footprint
        return take
    def cost(r):
        return r.cost if batch_cost is not None else r.tables.n_slots

    order = sorted(pending, key=lambda r: cost(r) /
                   (1.0 + max(0.0, now - r.t_arrival) / age_tau))
    overdue = [r for r in pending if now - r.t_arrival > max_wait]
    if overdue:
        order.remove(overdue[0])
        order.insert(0, overdue[0])
    spent = 0.0
    for r in order:
        if len(take) >= max_batch:
            break
        if take and batch_cost is None and r.tables.n_slots > 2 * take[0].tables.n_slots:
            continue
        if take and batch_cost is not None and spent + r.cost > max(take[0].cost, batch_cost):
            continue
        if need + r.footprint > budget_left:
            if not take:
                return []   # the best-scored request waits for memory
            continue
        take.append(r)
        need += r.footprint
        spent += cost(r)
    return take


def latency_percentiles(reqs, qs=(50, 99)):
    lat = np.array([r.latency() for r in reqs if r.done])
    return {f"p{q}": float(np.percentile(lat, q)) for q in qs} if lat.size else {}
