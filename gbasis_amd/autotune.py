"""Measured choice of kernel variants per work bucket.

For every bucket (class, primitive-count bin, precision class, dense flag,
sub-quartet phase) of a work queue the tuner times the admissible variants on the
bucket's own quartets and keeps the fastest:

* mixed precision: a bucket whose quartets may run with the FP32 recursion
  bulk (bound below the cutoff) is timed in FP32 and in FP64.  Both meet the
  error bound, so the faster one is kept.  (On CDNA4 non-packed FP32 FMA runs
  at the FP64 vector rate, so FP32 wins only where it relieves registers or LDS.)
* split classes: lanes per shell quartet R in {1, 2, 4, ..., 64} with either
  horizontal-recursion pass, or the cooperative kernel instead, or (6 <= L <= 17) the
  Rys primitive pass in either accumulator layout;
* classes of the cooperative kernel with 6 <= L <= 17: its own primitive pass or the
  Rys primitive pass in either layout (all finished by a horizontal-recursion pass);
* cooperative classes: primitive quartets in flight PF.

Choices are cached in memory per (class, bin, dense, precision class, phase) and can
be saved to / loaded from a JSON file.
"""

import json
import os

import torch

from ._lib import GBX_FP32_BULK, GBX_FP64


class Autotuner:
    FP32_MARGIN = 1.10
    # Rys pass variants: the two accumulator layouts (a third entry, when present, caps
    # the primitive quartets per LDS batch; the library's default is its largest batch)
    RYS = [("rys", 0), ("rys", 1)]

    def __init__(self, engine, repeats=3):
        self.eng = engine
        self.repeats = repeats
        self.choice = {}
        self.fp32_win = {}

    @staticmethod
    def key(bk):
        """(class, primitive bin, dense, precision, phase).  The phase keeps the small
        buckets of the build-centre sub-quartets (phases 1-3) from inheriting the choice
        measured on the much fuller phase-0 bucket of the same class and bin; like the
        other fields it is a property of the sub-quartet, so a request's variants do not
        depend on the batch it runs in."""
        return (tuple(bk.cls), int(bk.nprimq), bool(bk.dense), int(bk.precision),
                int(bk.phase))

    def _time(self, fn):
        s = torch.cuda.current_stream()
        fn()
        a = torch.cuda.Event(enable_timing=True)
        b = torch.cuda.Event(enable_timing=True)
        a.record(s)
        for _ in range(self.repeats):
            fn()
        b.record(s)
        torch.cuda.synchronize()
        return a.elapsed_time(b) / self.repeats

    def tune(self, dev, plans, out, log=None):
        """Pick variants for every bucket of ``plans`` (timed on ``out``).

        Classes whose FP32-eligible buckets run faster in FP64 lose the FP32
        permission in the engine's queue mask, so their quartets join the FP64
        launch from the next queue build on.
        """
        import time

        eng = self.eng
        fp32_win = dict(self.fp32_win)
        t_last = time.perf_counter()
        for ib, bk in enumerate(plans):
            if log is not None and time.perf_counter() - t_last > 20.0:
                log(f"  autotuner: {ib}/{len(plans)} buckets")
                t_last = time.perf_counter()
            k = self.key(bk)
            if k in self.choice:
                continue
            precisions = [GBX_FP64] if bk.precision == GBX_FP64 else [GBX_FP32_BULK, GBX_FP64]
            best = None
            for prec in precisions:
                variants = [None]
                rys = (eng.rys_ok(bk.cls) and dev.potential == "coulomb" and eng.path != "coop"
                       and prec == GBX_FP64)
                if not bk.dense and eng.has_reg(bk.cls):
                    variants = [None]
                elif eng.split_count(bk.cls) > 0 and eng.path != "coop":
                    hrrs = (0, 1) if (eng.has_hrr(bk.cls) and not bk.dense) else (0,)
                    variants = [("R", r, h) for r in (1, 2, 4, 8, 16, 32, 64) if r <= bk.nprimq
                                for h in hrrs] + [("coop",)] + (self.RYS if rys else [])
                elif rys:
                    variants = [("coop",)] + self.RYS
                for var in variants:
                    def run(prec=prec, var=var):
                        if var and var[0] == "rys":
                            eng.launch_class(dev, bk.cls, bk.quart, bk.nprimq, GBX_FP64,
                                             bk.mode(), out, bk.dense, "rys", seg=bk.seg,
                                             count=bk.count, rys_layout=var[1],
                                             rys_batch=var[2] if len(var) > 2 else None)
                            return
                        named = var[0] if var and var[0] == "coop" else None
                        eng.launch_bucket(dev, bk, out, precision=prec,
                                          lanes=var[1] if var and not named else None,
                                          hrr=bool(var[2]) if var and len(var) > 2 else None,
                                          path=named)
                    t = self._time(run)
                    if prec == GBX_FP32_BULK:
                        # keep FP32 only for a clear win: an FP64 bucket also merges with
                        # the class's other FP64 quartets into fewer, fuller launches
                        t *= self.FP32_MARGIN
                    if best is None or t < best[0]:
                        best = (t, prec, var)
            self.choice[k] = (best[1], best[2])
            if bk.precision == GBX_FP32_BULK:
                ck = (tuple(bk.cls), bool(bk.dense))
                fp32_win[ck] = fp32_win.get(ck, True) and best[1] == GBX_FP32_BULK
        if fp32_win != self.fp32_win:
            self.fp32_win = fp32_win
            eng.set_fp32_allowed(fp32_win)
        self.apply(plans)

    def apply(self, plans):
        for bk in plans:
            c = self.choice.get(self.key(bk))
            if c is not None:
                bk.run_precision, bk.variant = c

    def save(self, path):
        with open(path, "w") as fh:
            json.dump({"choice": [[list(k), v[0], list(v[1]) if v[1] else None]
                                  for k, v in self.choice.items()],
                       "fp32": [[list(k[0]), k[1], v] for k, v in self.fp32_win.items()]}, fh)

    def load(self, path):
        """Load saved choices; returns True if a cache was found."""
        if not os.path.exists(path):
            return False
        d = json.load(open(path))
        from .tables import NUM_PHASES

        for k, prec, var in d["choice"]:
            # choices saved without a phase apply to every phase
            phases = [k[4]] if len(k) > 4 else range(NUM_PHASES)
            for ph in phases:
                self.choice[(tuple(k[0]), k[1], k[2], k[3], ph)] = (prec,
                                                                   tuple(var) if var else None)
        self.fp32_win = {(tuple(c), dense): v for c, dense, v in d["fp32"]}
        self.eng.set_fp32_allowed(self.fp32_win)
        return True
