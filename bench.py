"""Benchmark of the ERI path on one MI355X.

Workload (default): 64 molecules of 4-24 atoms (seeds 1000..1063), per atom
1-3 s, 0-2 p, 0-1 d Cartesian shells, K = 1..6, exponents log-uniform in
[0.05, 1e3], Schwarz threshold 1e-12, canonical 8-fold packed FP64 output.
A step builds the screened, class-bucketed quartet queue on the device and
evaluates every unscreened shell quartet of the whole batch into the packed
array; the shell and pair tables and the Schwarz bounds are resident in HBM
before timing starts.  It runs twice: plain FP64, and mixed
(FP32 recursion bulk for quartets with Q_ij Q_kl below the cutoff; Boys,
prefactors and accumulation in FP64).  The headline value is the mixed form
at max|err| <= 1e-8 Eh against the FP64 result.

Prints one JSON line (see the keys below).  --gpus N > 1 is refused: this
build runs on one GPU.
"""

import argparse
import json
import os
import sys
import time

import numpy as np
from threadpoolctl import threadpool_limits

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "ERIs/s, mixed vs FP64 at max|err|<=1e-8 Eh; serving p50/p99 latency"


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def build_batch(n_mol):
    from gbasis_amd.tables import BatchTables, MoleculeTables
    from gbasis_amd.workloads import batch_molecule

    mols = [MoleculeTables(batch_molecule(1000 + m)) for m in range(n_mol)]
    return BatchTables(mols)


def time_steps(eng, dev, bound, threshold, cutoff, steps, warmup, tuner=None):
    """warmup + steps of: device queue build, then every ERI launch (the launches read
    their quartet counts on the device: no host round trip inside a step).

    Returns (seconds per step, queue seconds per step, wall seconds per step,
    per-bucket kernel seconds per step, plans of the last step).
    """
    import torch

    stream = torch.cuda.current_stream()
    for w in range(warmup):
        t0 = time.perf_counter()
        plans = eng.plan_device(dev, bound, threshold, cutoff, stats=False, device_counts=True)
        if tuner is not None:
            tuner.tune(dev, plans, dev.out, log=log)
        eng.run(dev, plans, dev.out)
        torch.cuda.synchronize()
        log(f"  warm-up step {w + 1}/{warmup}: {time.perf_counter() - t0:.1f} s")
    per_bucket = None
    t_queue = 0.0
    t0 = time.perf_counter()
    start = torch.cuda.Event(enable_timing=True)
    stop = torch.cuda.Event(enable_timing=True)
    start.record(stream)
    recs = []
    for _ in range(steps):
        q0 = torch.cuda.Event(enable_timing=True)
        q0.record(stream)
        plans = eng.plan_device(dev, bound, threshold, cutoff, stats=False, device_counts=True)
        if tuner is not None:
            tuner.apply(plans)
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(len(plans) + 1)]
        marks[0].record(stream)
        if eng.launch_streams > 1:
            eng.run(dev, plans, dev.out)
            for m in marks[1:]:
                m.record(stream)
        else:
            for i, bk in enumerate(plans):
                bk.path = eng.launch_bucket(dev, bk, dev.out)
                marks[i + 1].record(stream)
        recs.append((q0, marks))
        if time.perf_counter() - t0 > 20.0:
            torch.cuda.synchronize()
            log(f"  step {len(recs)}/{steps} done after {time.perf_counter() - t0:.0f} s")
    stop.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    eng.check_queue(dev)   # no bucket overflowed (its quartets would be missing)
    per_bucket = np.zeros(len(plans))
    for q0, marks in recs:
        t_queue += q0.elapsed_time(marks[0]) * 1e-3
        for i in range(len(plans)):
            per_bucket[i] += marks[i].elapsed_time(marks[i + 1]) * 1e-3
    per_bucket /= steps
    if eng.launch_streams > 1:
        # per-class times from one more step with the launches in sequence (not in the total)
        marks = [torch.cuda.Event(enable_timing=True) for _ in range(len(plans) + 1)]
        marks[0].record(stream)
        for i, bk in enumerate(plans):
            eng.launch_bucket(dev, bk, dev.out)
            marks[i + 1].record(stream)
        torch.cuda.synchronize()
        per_bucket = np.array([marks[i].elapsed_time(marks[i + 1]) * 1e-3
                               for i in range(len(plans))])
    return (start.elapsed_time(stop) * 1e-3 / steps, t_queue / steps, wall / steps,
            per_bucket, plans)


def time_graph(eng, dev, bound, threshold, cutoff, steps, tuner=None):
    """The same step recorded once into a HIP graph and replayed ``steps`` times.
    Returns (seconds per replay, the output of the last replay)."""
    import torch

    graph, out, _ = eng.capture(dev, bound, threshold, cutoff, out=dev.out, tuner=tuner)
    graph.replay()
    torch.cuda.synchronize()
    a = torch.cuda.Event(enable_timing=True)
    b = torch.cuda.Event(enable_timing=True)
    a.record()
    for _ in range(steps):
        graph.replay()
    b.record()
    torch.cuda.synchronize()
    return a.elapsed_time(b) * 1e-3 / steps, out


def _host_busy(srv, span):
    """Fraction of the wall span the serving thread spent building tables / enqueueing."""
    return {k: v / span for k, v in srv.host_seconds.items()}


def serving_bench(n_requests, n_sat, threshold, cutoff, n_streams, seed=4, max_batch=32,
                  batch_flops=8e12):
    """Serving workload: requests of 2-30 atoms (seeds 4000 + i, s/p/d, K <= 6),
    precision FP64 or mixed with probability 1/2 each.  First the saturated rate
    (n_sat requests all arriving at once), then Poisson arrivals at 50 % and 80 %
    of it with n_requests each; latency = completion (packed result resident in
    HBM) minus scheduled arrival, which includes the host table build."""
    import torch
    from gbasis_amd.engine import EriEngine
    from gbasis_amd.serving import EriServer, latency_percentiles
    from gbasis_amd.tables import BatchTables
    from gbasis_amd.engine import DeviceTables
    from gbasis_amd.workloads import serving_request

    rng = np.random.default_rng(seed)
    bases = [serving_request(4000 + i) for i in range(n_requests)]
    precs = np.where(rng.random(n_requests) < 0.5, "mixed", "fp64")
    srv = EriServer(threshold=threshold, fp32_cutoff=cutoff, n_streams=n_streams,
                    max_batch=max_batch, batch_flops=batch_flops)
    clock = srv.clock
    # warm-up: programs and per-stream workspaces at the sizes of full batches (a
    # workspace that grows inside a timed run costs that batch a device allocation)
    for i in range(min(n_requests, 2 * max_batch * n_streams)):
        srv.submit(bases[i], precs[i])
    srv.drain(on_done=srv.release)

    def run(count, rate):
        arr = (np.cumsum(rng.exponential(1.0 / rate, count)) if rate else np.zeros(count))
        reqs = []
        t0 = clock()
        i = ndone = 0
        last_progress = last_log = t0
        while i < count or srv.busy():
            now = clock() - t0
            while i < count and arr[i] <= now:
                reqs.append(srv.submit(bases[i], str(precs[i]), t_arrival=t0 + arr[i]))
                i += 1
            done = srv.step()
            for r in done:
                srv.release(r)
            ndone += len(done)
            t = clock()
            if done:
                last_progress = t
            elif t - last_progress > 60.0:
                raise RuntimeError(f"serving made no progress for 60 s ({ndone}/{count} done)")
            else:
                time.sleep(20e-6)
            if t - last_log > 10.0:
                log(f"  serving: {ndone}/{count} requests done after {t - t0:.0f} s")
                last_log = t
        span = max(r.t_done for r in reqs) - t0
        return reqs, span

    def eris_of(reqs):
        eng = EriEngine()
        total = 0
        mols = [r.tables for r in reqs]
        for c in range(0, len(mols), 64):
            dev = DeviceTables(BatchTables(mols[c:c + 64]))
            bound = eng.pair_bounds(dev, eng.schwarz(dev))
            total += sum(bk.eris for bk in eng.plan_device(dev, bound, threshold, 0.0))
        torch.cuda.synchronize()
        return total

    srv.host_seconds = {"tables": 0.0, "launch": 0.0}
    sat_reqs, sat_span = run(n_sat, None)
    sat_rate = n_sat / sat_span
    sat_eris = eris_of(sat_reqs)
    hb = _host_busy(srv, sat_span)
    log(f"serving saturated: {sat_rate:.1f} req/s, {sat_eris / sat_span:.3e} ERIs/s "
        f"({n_sat} requests, {srv.batches_launched} batches so far); host busy "
        f"{hb['tables']:.0%} tables + {hb['launch']:.0%} launch")
    out = {"workload": "requests of 2-30 atoms (s/p/d Cartesian, K 1-6, exps [0.05,1e3]), "
                       "FP64 or mixed with probability 1/2, Schwarz threshold "
                       f"{threshold:g}; result = packed ERIs resident in HBM",
           "streams": n_streams,
           "batching": {"max_requests": max_batch, "max_flops": batch_flops,
                        "rule": "summed estimated FLOPs of a batch <= max(opener's, max_flops)"},
           "saturated": {"requests": n_sat, "req_per_s": sat_rate,
                                               "eris_per_s": sat_eris / sat_span}}
    all_eris = None
    out["saturated"]["host_busy"] = _host_busy(srv, sat_span)
    for frac in (0.5, 0.8):
        b0 = srv.batches_launched
        srv.host_seconds = {"tables": 0.0, "launch": 0.0}
        reqs, span = run(n_requests, frac * sat_rate)
        if all_eris is None:
            all_eris = eris_of(reqs)
        lat = latency_percentiles(reqs)
        res = {"lambda_req_per_s": frac * sat_rate, "requests": n_requests,
               "p50_ms": lat["p50"] * 1e3, "p99_ms": lat["p99"] * 1e3,
               "sustained_req_per_s": n_requests / span,
               "sustained_eris_per_s": all_eris / span,
               "batches": srv.batches_launched - b0, "host_busy": _host_busy(srv, span)}
        log(f"serving at {frac:.0%} ({frac * sat_rate:.1f} req/s): p50 {res['p50_ms']:.1f} ms, "
            f"p99 {res['p99_ms']:.1f} ms, {res['sustained_eris_per_s']:.3e} ERIs/s; host busy "
            f"{res['host_busy']['tables']:.0%} tables + {res['host_busy']['launch']:.0%} launch")
        out[f"at_{int(frac * 100)}pct"] = res
    del srv
    torch.cuda.empty_cache()
    return out


def attach_stats(timed, stats_plans):
    """Copy quartet / ERI / FLOP counts of the counting pass onto the timed buckets
    (matched by bucket key; a timed bucket with no counterpart held no quartets)."""
    by_key = {bk.key: bk for bk in stats_plans}
    for bk in timed:
        s = by_key.get(bk.key)
        bk.n, bk.eris, bk.flops = (s.n, s.eris, s.flops) if s is not None else (0, 0, 0.0)
    return timed


def class_table(plans, per_bucket, top=16):
    """Per (kernel path, class, precision): time, quartets, ERIs, algorithmic TFLOP/s."""
    groups = {}
    for bk, t in zip(plans, per_bucket):
        key = f"{bk.path}:{''.join(map(str, bk.cls))}:{'f32' if bk.launch_precision() else 'f64'}"
        g = groups.setdefault(key, [0.0, 0, 0, 0.0])
        g[0] += t
        g[1] += bk.n
        g[2] += bk.eris
        g[3] += bk.flops
    rows = sorted(groups.items(), key=lambda kv: -kv[1][0])[:top]
    return [{"kernel": k, "ms": v[0] * 1e3, "quartets": v[1], "eris": v[2],
             "tflops": v[3] / v[0] / 1e12 if v[0] > 0 else 0.0} for k, v in rows]


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "eri_batch64_hbm_counters.json")


def recorded_traffic(kernel):
    """HBM bytes per launch of ``kernel`` from the recorded FETCH_SIZE / WRITE_SIZE runs
    (each scaled by the counter's reading for a plain copy of known size), or None."""
    if not os.path.exists(TRAFFIC_FILE):
        return None, None
    kernels = json.load(open(TRAFFIC_FILE)).get("kernels", {})
    rec = kernels.get(kernel, kernels.get("gbx::" + kernel))
    if rec is None:
        return None, None
    return rec["fetch_bytes"] + rec["write_bytes"], os.path.relpath(TRAFFIC_FILE, ROOT)


def roofline_entry(plans, per_bucket, prefer_reg=True):
    """Roofline of the dominant kernel: the fused register kernel (one launch per step)
    with the largest time (any kernel when ``prefer_reg`` is off or none ran);
    algorithmic FLOPs per launch / its average launch time, against the vector peak
    of its precision.  Also the aggregate of all ERI kernels."""
    from gbasis_amd.roofline import PEAK_FP32_VECTOR, PEAK_FP64_VECTOR

    groups = {}
    for bk, t in zip(plans, per_bucket):
        key = (bk.path, bk.cls, bk.launch_precision())
        g = groups.setdefault(key, [0.0, 0.0, 0])
        g[0] += t
        g[1] += bk.flops
        g[2] += 1
    fused = ({k: v for k, v in groups.items() if k[0] == "reg"} if prefer_reg else {}) or groups
    key, (t, f, nl) = max(fused.items(), key=lambda kv: kv[1][0])
    path, cls, prec = key
    peak = PEAK_FP64_VECTOR if prec == 0 else PEAK_FP32_VECTOR
    ach = f / t
    name = f"eri_{path}_{''.join(map(str, cls))}<{'double' if prec == 0 else 'float'}>"
    total_t = float(per_bucket.sum())
    total_f = float(sum(bk.flops for bk in plans))
    traffic, tsrc = recorded_traffic(name)
    # compute-bound ("mfma" in the hbm / mfma vocabulary of the bench line): the
    # recursions issue FP64 vector FMAs, whose 78.6 TFLOP/s peak equals the FP64 MFMA peak
    return {"bound": "mfma", "compute_unit": "fp64 vector ALU (FMA chains, no MFMA); same "
            "78.6 TFLOP/s peak as fp64 MFMA", "kernel": name, "achieved": ach / 1e12,
            "peak": peak / 1e12,
            "unit": "TFLOP/s", "frac": ach / peak, "traffic": traffic, "traffic_source": tsrc,
            "launches_per_step": nl,
            "flops_per_launch": f / nl, "avg_launch_ms": t / nl * 1e3,
            "kernel_time_share": t / total_t,
            "all_kernels": {"achieved": total_f / total_t / 1e12, "frac": total_f / total_t / peak}}


def time_schwarz(eng, dev, reps=3):
    """Milliseconds of the Schwarz pass (the diagonal (ij|ij) quartets of every shell pair
    and the pair bounds), warm, outside the timed step."""
    import torch

    eng.pair_bounds(dev, eng.schwarz(dev))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        eng.pair_bounds(dev, eng.schwarz(dev))
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps * 1e3


def fp32_fraction(plans_mixed):
    """Fraction of the shell quartets whose recursion bulk actually runs in FP32 in the
    mixed form (after the autotuner's choices)."""
    tot = sum(bk.n for bk in plans_mixed if bk.stores())
    f32 = sum(bk.n for bk in plans_mixed if bk.stores() and bk.launch_precision() == 1)
    return f32 / tot if tot else 0.0


def mixed_detail(plans_mixed, t64, tmx):
    """What the mixed form is, in numbers."""
    return (f"f64 Boys, prefactors, accumulation and HRR; f32 recursion bulk for "
            f"{100 * fp32_fraction(plans_mixed):.2f} % of the shell quartets (those below the "
            f"f32 cutoff that the autotuner measured faster in f32); mixed step / f64 step = "
            f"{tmx / t64:.3f}: on this chip non-packed f32 issues at the f64 rate, so the "
            f"mixed form is an f64 computation in all but those quartets")


def cpu_baselines(batch, plans_fp64, budget_s=12.0):
    """The oracle's CPU rate on a stratified sample of the step's shell quartets: timed
    by one worker process per host core (at most 16, BLAS on one thread each, side by
    side), and in this process on one core.  Returns (all cores, one core)."""
    stats, total_eris = baseline_sample(batch, plans_fp64)
    ncores = max(1, min(16, len(os.sched_getaffinity(0))))
    with threadpool_limits(limits=1):
        one = baseline_one_core(stats, total_eris, budget_s)
    try:
        allc = baseline_all_cores(stats, total_eris, budget_s, ncores)
    except (OSError, RuntimeError, ValueError) as exc:
        print(f"[bench] all-core CPU baseline failed ({exc}); reporting the one-core one",
              file=sys.stderr, flush=True)
        allc = one
    return allc, one


def baseline_sample(batch, plans_fp64, per_class=256):
    """Up to ``per_class`` random shell quartets of every canonical class of the step
    (drawn from the storing buckets: one sub-quartet per shell quartet, mapped back to the
    shell pairs), as four shells [l, coord, exps, coeffs] each; classes sorted by their
    algorithmic FLOPs per quartet.  Returns ([(class, quartets in the step, samples)],
    ERIs of the step)."""
    from gbasis_amd.programs import canonical_class

    rng = np.random.default_rng(0)
    pair_mol = np.zeros(batch.npair, dtype=np.int64)
    for mi, m in enumerate(batch.mols):
        pair_mol[batch.mol_pair_off[mi]:batch.mol_pair_off[mi] + m.npair] = mi
    by_cls, flops_cls = {}, {}
    for bk in plans_fp64:
        cc = canonical_class(*bk.cls)[0]
        flops_cls[cc] = flops_cls.get(cc, 0.0) + bk.flops
        if bk.n > 0 and bk.stores():
            by_cls.setdefault(cc, []).append(bk)
    stats = []
    for cls, bks in by_cls.items():
        n = sum(bk.n for bk in bks)
        quart = batch.entry_pair[np.concatenate([bk.host_quartets() for bk in bks])]
        samples = []
        for q in quart[rng.integers(0, quart.shape[0], size=min(per_class, quart.shape[0]))]:
            mi = pair_mol[q[0]]
            m = batch.mols[mi]
            po = batch.mol_pair_off[mi]
            sh = [m.pair_shell_a[q[0] - po], m.pair_shell_b[q[0] - po],
                  m.pair_shell_a[q[1] - po], m.pair_shell_b[q[1] - po]]
            samples.append([[int(m.shells[i].l), m.shells[i].coord.tolist(),
                             m.shells[i].exps.tolist(), m.shells[i].coeffs.tolist()]
                            for i in sh])
        stats.append((flops_cls[cls] / n, cls, n, samples))
    stats.sort(key=lambda t: t[0])
    return [(cls, n, samples) for _, cls, n, samples in stats], \
        sum(bk.eris for bk in plans_fp64)


_BASELINE_NOTE = ("; oracle = NumPy restatement of the reference OS+HGP engine, timed per "
                  "class on random quartets of the step and extrapolated by class counts")


def baseline_one_core(stats, total_eris, budget_s):
    """The sample timed in this process (BLAS limited to one thread by the caller)."""
    from oracle.rate import time_classes

    times = time_classes([s for _, _, s in stats], 0, budget_s)
    est = sum(n * t / k for (_, n, _), (t, k) in zip(stats, times))
    spent = sum(t for t, _ in times)
    return {"value": total_eris / est, "unit": "ERIs/s", "cores": 1, "kind": "port",
            "sample": f"{sum(k for _, k in times)} shell quartets over {len(stats)} classes "
                      f"({spent:.1f} s of CPU) in one process on one core" + _BASELINE_NOTE}


def baseline_all_cores(stats, total_eris, budget_s, ncores):
    """The sample timed by ``ncores`` worker processes at once (python -m oracle.rate),
    each on one BLAS thread with its own random draws: the per-class mean time under full
    load gives the one-core time of the step, divided by ``ncores``."""
    import subprocess
    import tempfile

    with tempfile.NamedTemporaryFile("w", suffix=".json", delete=False) as f:
        json.dump({"classes": [s for _, _, s in stats]}, f)
        path = f.name
    env = dict(os.environ, OMP_NUM_THREADS="1", OPENBLAS_NUM_THREADS="1", MKL_NUM_THREADS="1")
    try:
        procs = [subprocess.Popen([sys.executable, "-m", "oracle.rate", path, str(1 + w),
                                   str(budget_s)], cwd=ROOT, env=env, stdout=subprocess.PIPE,
                                  stderr=subprocess.PIPE, text=True) for w in range(ncores)]
        results = []
        for p in procs:
            out, err = p.communicate(timeout=budget_s * 4 + 300)
            if p.returncode != 0:
                raise RuntimeError(f"oracle.rate worker exited {p.returncode}: {err[-300:]}")
            results.append(json.loads(out.strip().splitlines()[-1]))
    finally:
        os.unlink(path)
    tsum = np.sum([[t for t, _ in r] for r in results], axis=0)
    ksum = np.sum([[k for _, k in r] for r in results], axis=0)
    est = sum(n * t / k for (_, n, _), t, k in zip(stats, tsum, ksum))
    return {"value": float(total_eris * ncores / est), "unit": "ERIs/s", "cores": ncores,
            "kind": "port",
            "sample": f"{int(ksum.sum())} shell quartets over {len(stats)} classes "
                      f"({float(tsum.sum()):.1f} s of CPU) in {ncores} worker processes at "
                      "once, one core each" + _BASELINE_NOTE}


CONFIG_WORKLOADS = {
    # name: (description, Schwarz threshold, dense output)
    "dense_small": ("4 atoms in a 4 Bohr cube, Cartesian s,s,p,d per atom (N=44), K=3, exps "
                    "[0.1,50], no screening, dense chemist output", 0.0, True),
    "large_spdf": ("16 atoms in a 10 Bohr cube, Cartesian s,p,d,f per atom (N=320), K=4, exps "
                   "[0.08,2e3], Schwarz 1e-10, packed output", 1e-10, False),
    "high_l_pure": ("6 atoms in a 6 Bohr cube, pure s,p,d,f,g per atom (N=150), K=8, exps "
                    "[0.02,1e5], Schwarz 1e-12, packed output", 1e-12, False),
}
SAMPLES_FILE = os.path.join(ROOT, "tests", "golden", "config_samples.npz")
# workloads whose packed ERIs also go through the four-index transform (square, seeded
# transform): the dense new-basis array is nfunc^4 doubles (4 GB at N = 150)
MO_WORKLOADS = ("high_l_pure",)


def time_mo_transform(eng, packed, nfunc, reps=3):
    """Milliseconds and FP64 rate of the four-index transform of the packed ERIs into a
    basis of nfunc functions (gbx_mo_transform: two passes of Y = T X T^T over n x n
    matrices, n^2 m + n m^2 / 2 algorithmic FMAs per matrix, on the FP64 matrix cores,
    whose MI355X peak equals the FP64 vector peak of 78.6 TFLOP/s; npair(n) matrices in
    the first pass, npair(m) in the second)."""
    import torch

    from gbasis_amd.roofline import PEAK_FP64_VECTOR

    rng = np.random.default_rng(11)
    T = rng.standard_normal((nfunc, nfunc)) / np.sqrt(nfunc)
    out = eng.mo_transform(packed, nfunc, T, "chemist")
    del out
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    for _ in range(reps):
        out = eng.mo_transform(packed, nfunc, T, "chemist")
        del out
    ev[1].record()
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1]) / reps
    n = m = nfunc
    npn, npm = n * (n + 1) // 2, m * (m + 1) // 2
    flops = 2.0 * (npn + npm) * (n * n * m + n * m * m / 2)
    log(f"  four-index transform N={n}: {ms:.2f} ms, {flops / ms / 1e9:.2f} TFLOP/s")
    return {"n": n, "m": m, "ms": ms, "flops": flops, "tflops": flops / ms / 1e9,
            "frac": flops / ms / 1e9 / (PEAK_FP64_VECTOR / 1e12),
            "includes": "workspace and output allocation (torch caching allocator)"}


TUNE_FILE = os.path.join(ROOT, "gbasis_amd", "tuning_mi355x.json")


def _packed_index(i, j, k, l):
    i, j = np.maximum(i, j), np.minimum(i, j)
    k, l = np.maximum(k, l), np.minimum(k, l)
    ij = i * (i + 1) // 2 + j
    kl = k * (k + 1) // 2 + l
    ij, kl = np.maximum(ij, kl), np.minimum(ij, kl)
    return ij * (ij + 1) // 2 + kl


def _canon_cls(ls):
    la, lb, lc, ld = ls
    a, b = max(la, lb), min(la, lb)
    c, d = max(lc, ld), min(lc, ld)
    return max((a, b, c, d), (c, d, a, b))


def sample_errors(name, mol, basis, q_orig, threshold, outs):
    """Per-class errors of the GPU outputs ``outs`` {label: packed device array} at the
    oracle's sampled quartets (tests/golden/config_samples.npz), against the
    binary128 oracle; also the reference algorithm's own FP64 error there."""
    import torch

    if not os.path.exists(SAMPLES_FILE):
        return None
    z = np.load(SAMPLES_FILE)
    if f"{name}_idx" not in z:
        return None
    idx, shapes = z[f"{name}_idx"], z[f"{name}_shape"]
    exact, ref64 = z[f"{name}_quad"], z[f"{name}_fp64"]
    first = {}
    for es in mol.shells:
        first.setdefault(es.orig, es.func_off)
    rows = {}
    pos = 0
    screened = 0
    for q, shp in zip(idx.tolist(), shapes.tolist()):
        n = int(np.prod(shp))
        ex, rf = exact[pos:pos + n], ref64[pos:pos + n]
        pos += n
        if threshold > 0 and q_orig[q[0], q[1]] * q_orig[q[2], q[3]] < threshold:
            screened += 1
            continue
        a, b, c, d = np.meshgrid(*[first[x] + np.arange(m) for x, m in zip(q, shp)],
                                 indexing="ij")
        slots = torch.from_numpy(_packed_index(a, b, c, d).ravel()).to("cuda")
        cls = "".join(map(str, _canon_cls([basis[x].angmom for x in q])))
        r = rows.setdefault(cls, {"quartets": 0, "max_abs": 0.0,
                                  "ref_fp64_err": 0.0,
                                  **{f"{k}_err": 0.0 for k in outs}})
        r["quartets"] += 1
        r["max_abs"] = max(r["max_abs"], float(np.abs(ex).max()))
        r["ref_fp64_err"] = max(r["ref_fp64_err"], float(np.abs(rf - ex).max()))
        for k, out in outs.items():
            got = out[slots].cpu().numpy()
            r[f"{k}_err"] = max(r[f"{k}_err"], float(np.abs(got - ex).max()))
    summary = {"sampled_quartets": int(idx.shape[0]), "screened_on_gpu": screened,
               "classes": len(rows)}
    for k in list(outs) + ["ref_fp64"]:
        summary[f"max_{k}_err"] = max((r[f"{k}_err"] for r in rows.values()), default=0.0)
    return {"summary": summary, "per_class": rows}


def per_class_mixed_error(eng, dev, plans64, plansmx, n_slots):
    """Max / RMS of mixed - FP64 over the packed slots of each class (each class run
    alone into zeroed buffers)."""
    import torch

    from gbasis_amd.programs import canonical_class

    def canon(bk):
        return canonical_class(*bk.cls)[0]

    a = torch.zeros(n_slots, dtype=torch.float64, device="cuda")
    b = torch.zeros_like(a)
    out = {}
    fp32_classes = {canon(bk) for bk in plansmx if bk.launch_precision() == 1 and bk.n > 0}
    for cls in sorted({canon(bk) for bk in plans64}):
        if cls not in fp32_classes:
            # no quartet of the class runs the FP32 bulk: mixed == FP64 bit for bit
            out["".join(map(str, cls))] = {"slots": None, "max_abs_err": 0.0, "rms_err": 0.0,
                                           "fp32_quartets": 0}
            continue
        a.zero_()
        b.zero_()
        # every phase of the class's shell quartets, in phase order
        for bk in plans64:
            if canon(bk) == cls:
                eng.launch_bucket(dev, bk, a)
        for bk in plansmx:
            if canon(bk) == cls:
                eng.launch_bucket(dev, bk, b)
        nz = a != 0
        d = (b - a).abs_()
        cnt = int(nz.sum())
        out["".join(map(str, cls))] = {
            "slots": cnt, "max_abs_err": float(d.max()),
            "rms_err": float(torch.sqrt((d[nz] ** 2).mean())) if cnt else 0.0,
            "fp32_quartets": sum(bk.n for bk in plansmx
                                 if canon(bk) == cls and bk.stores()
                                 and bk.launch_precision() == 1)}
    del a, b
    return out


def config_bench(args):
    """One single-molecule workload (``--workload NAME``): prints its JSON line."""
    print(json.dumps(run_workload(args.workload, args)), flush=True)


def workload_summary(r):
    """The keys of one single-molecule workload's line that the default bench line
    carries for it."""
    s = r.get("oracle_samples") or {}
    summ = s.get("summary", {})
    per = s.get("per_class", {})
    return {
        "workload": r["config"]["workload"],
        "eris": r["config"]["eris"], "shell_quartets": r["config"]["shell_quartets"],
        "fp64": r["fp64"], "mixed": {"value": r["value"], "ms_per_step": r["ms_per_step"]},
        "step_timed_as": r["step_timed_as"],
        "schwarz_ms": r.get("schwarz_ms"),
        "dense_unpack_ms": r.get("dense_unpack_ms"),
        "mixed_vs_fp64": {k: r["mixed_vs_fp64"][k] for k in ("max_abs_err", "rms_err")},
        "fp32_quartet_fraction": r.get("fp32_quartet_fraction"),
        "vs_binary128": {
            "sampled_quartets": summ.get("sampled_quartets"),
            "max_gpu_fp64_err": summ.get("max_gpu_fp64_err"),
            "max_gpu_mixed_err": summ.get("max_gpu_mixed_err"),
            "max_ref_fp64_err": summ.get("max_ref_fp64_err"),
            "per_class_gpu_fp64_err": {k: v["gpu_fp64_err"] for k, v in per.items()},
        },
        "roofline": r["roofline"],
        "classes_fp64": r["classes_fp64"][:6],
    }


def make_tuner(eng, args):
    """The variant autotuner with the choices of ``--tune-cache`` loaded (by default the
    MI355X choices shipped in gbasis_amd/tuning_mi355x.json; buckets it does not list
    are timed when first seen), or None with --no-tune."""
    from gbasis_amd.autotune import Autotuner

    if args.no_tune:
        return None
    tuner = Autotuner(eng)
    if args.tune_cache and tuner.load(args.tune_cache):
        log(f"autotuner choices loaded from {args.tune_cache} ({len(tuner.choice)} buckets)")
    return tuner


def run_workload(name, args, tuner=None):
    """One single-molecule workload: FP64 and mixed steps, per-class errors, roofline,
    CPU baseline.  Returns its JSON record.  ``tuner``: shared autotuner (default: a new
    one from make_tuner)."""
    import torch
    from gbasis_amd import workloads
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables

    desc, thr, dense = CONFIG_WORKLOADS[name]
    basis = getattr(workloads, name)()
    mol = MoleculeTables(basis)
    batch = BatchTables([mol])
    eng = get_engine("cuda")
    eng.launch_streams = args.launch_streams
    eng.hrr_lds = args.hrr_lds
    dev = DeviceTables(batch)
    schwarz_ms = time_schwarz(eng, dev)
    q = eng.schwarz(dev)
    bound = eng.pair_bounds(dev, q)
    q_orig = mol.schwarz_orig(q[:, 0].cpu().numpy())
    plans64 = eng.plan_device(dev, bound, thr, 0.0, stats=True)
    eris = sum(bk.eris for bk in plans64)
    nquart = sum(bk.n for bk in plans64 if bk.stores())
    nsub = sum(bk.n for bk in plans64)
    log(f"{name}: N={mol.nfunc}, {nquart} shell quartets ({nsub} sub-quartets), {eris} ERIs, "
        f"{mol.n_slots} slots")
    if tuner is None:
        tuner = make_tuner(eng, args)
    dev.out = torch.zeros(batch.n_slots, dtype=torch.float64, device="cuda")
    t64, q64, _, pb64, last64 = time_steps(eng, dev, bound, thr, 0.0, args.steps, args.warmup,
                                           tuner)
    g64 = None if args.no_graph else time_graph(eng, dev, bound, thr, 0.0, args.steps, tuner)[0]
    out64 = dev.out
    log(f"{name} fp64: {t64 * 1e3:.2f} ms/step, {eris / t64:.3e} ERIs/s"
        + ("" if g64 is None else f"; graph {g64 * 1e3:.2f} ms/step"))
    unpack_ms = None
    if dense:
        ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
        eng.unpack_dense(out64, mol.nfunc, "chemist")
        ev[0].record()
        for _ in range(args.steps):
            d = eng.unpack_dense(out64, mol.nfunc, "chemist")
        ev[1].record()
        torch.cuda.synchronize()
        unpack_ms = ev[0].elapsed_time(ev[1]) / args.steps
        del d
    mo = time_mo_transform(eng, out64, mol.nfunc) if name in MO_WORKLOADS else None
    dev.out = torch.zeros(batch.n_slots, dtype=torch.float64, device="cuda")
    tmx, qmx, _, pbmx, lastmx = time_steps(eng, dev, bound, thr, args.fp32_cutoff, args.steps,
                                           args.warmup, tuner)
    gmx = (None if args.no_graph else
           time_graph(eng, dev, bound, thr, args.fp32_cutoff, args.steps, tuner)[0])
    outmx = dev.out
    log(f"{name} mixed: {tmx * 1e3:.2f} ms/step, {eris / tmx:.3e} ERIs/s"
        + ("" if gmx is None else f"; graph {gmx * 1e3:.2f} ms/step"))
    plansmx = eng.plan_device(dev, bound, thr, args.fp32_cutoff, stats=True)
    if tuner is not None:
        tuner.apply(plansmx)
        tuner.apply(plans64)
    attach_stats(last64, plans64)
    attach_stats(lastmx, plansmx)
    diff = (outmx - out64).abs_()
    max_err = float(diff.max())
    rms_err = float(torch.linalg.vector_norm(diff)) / np.sqrt(diff.numel())
    del diff
    samples = sample_errors(name, mol, basis, q_orig, thr, {"gpu_fp64": out64, "gpu_mixed": outmx})
    per_cls = per_class_mixed_error(eng, dev, plans64, plansmx, batch.n_slots)
    t64h = g64 if g64 is not None else t64
    tmxh = gmx if gmx is not None else tmx
    result = {
        "metric": METRIC, "value": eris / tmxh, "unit": "ERIs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": tmxh * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32",
        "dtype_detail": mixed_detail(plansmx, t64h, tmxh),
        "data": "synthetic",
        "config": {"workload": desc, "name": name, "nfunc": mol.nfunc, "shell_quartets": nquart,
                   "sub_quartets": nsub,
                   "eris": eris, "packed_slots": mol.n_slots, "schwarz_threshold": thr,
                   "fp32_cutoff": args.fp32_cutoff},
        "step_timed_as": "hip graph replay" if gmx is not None else "eager launches",
        "fp64": {"value": eris / t64h, "ms_per_step": t64h * 1e3},
        "eager_ms_per_step": {"fp64": t64 * 1e3, "mixed": tmx * 1e3},
        "schwarz_ms": schwarz_ms,
        "fp32_quartet_fraction": fp32_fraction(plansmx),
        "dense_unpack_ms": unpack_ms,
        "mo_transform": mo,
        "mixed_vs_fp64": {"max_abs_err": max_err, "rms_err": rms_err, "per_class": per_cls},
        "oracle_samples": samples,
        "roofline": roofline_entry(last64, pb64, prefer_reg=False),
        "roofline_mixed": roofline_entry(lastmx, pbmx, prefer_reg=False),
        "classes_fp64": class_table(last64, pb64),
        "classes_mixed": class_table(lastmx, pbmx),
    }
    if not args.no_cpu:
        result["cpu_baseline"], result["cpu_baseline_one_core"] = cpu_baselines(batch, plans64)
    return result


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--molecules", type=int, default=64)
    ap.add_argument("--threshold", type=float, default=1e-12)
    ap.add_argument("--fp32-cutoff", type=float, default=1e-2)
    ap.add_argument("--path", choices=["auto", "coop"], default="auto")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="skip the variant autotuner")
    ap.add_argument("--tune-cache", default=TUNE_FILE,
                    help="autotuner choices to load (default: the shipped MI355X choices); "
                         "'' to tune every bucket afresh")
    ap.add_argument("--tune-save", default="",
                    help="write the autotuner's choices (loaded and new) here at the end")
    ap.add_argument("--copy-calibration-mb", type=int, default=0,
                    help="run one plain device copy of this size first (HBM counter calibration)")
    ap.add_argument("--only", choices=["both", "fp64"], default="both",
                    help="fp64: time the FP64 form only (for single-precision kernel traces)")
    ap.add_argument("--no-graph", action="store_true", help="skip the HIP-graph replay timing")
    ap.add_argument("--serving-requests", type=int, default=600)
    ap.add_argument("--serving-saturation", type=int, default=300)
    ap.add_argument("--serving-streams", type=int, default=4)
    ap.add_argument("--serving-max-batch", type=int, default=32)
    ap.add_argument("--serving-batch-flops", type=float, default=8e12,
                    help="cap on a serving batch's summed algorithmic FLOPs (0: size batches "
                         "by count and the 2x packed-slot rule only)")
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-workloads", action="store_true",
                    help="skip the single-molecule workloads in the default line")
    ap.add_argument("--workload-steps", type=int, default=2,
                    help="timed steps of each single-molecule workload in the default line")
    ap.add_argument("--launch-streams", type=int, default=4,
                    help="spread the class launches of a step over this many streams (the "
                         "per-class timings then need --launch-streams 1)")
    ap.add_argument("--hrr-lds", choices=["auto", "always", "never"], default="auto",
                    help="horizontal-recursion pass in LDS: for the classes whose cooperative "
                         "pass needs global scratch (auto), wherever it fits, or never")
    ap.add_argument("--workload", choices=["batch64"] + list(CONFIG_WORKLOADS), default="batch64",
                    help="batch64: the 64-molecule batch plus serving (default); the others "
                         "are single-molecule workloads")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: gbasis_amd measures the ERI path on one GPU only (multi-GPU "
              "sharding is a design, not an implementation); --gpus must be 1",
              file=sys.stderr)
        sys.exit(2)

    import torch
    from gbasis_amd.engine import DeviceTables, get_engine

    if args.workload != "batch64":
        config_bench(args)
        return
    if args.copy_calibration_mb > 0:
        n = args.copy_calibration_mb * (1 << 20) // 8
        src = torch.ones(n, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {n * 8} bytes read, {n * 8} bytes written")
        del src, dst
    t0 = time.perf_counter()
    batch = build_batch(args.molecules)
    log(f"host tables: {len(batch.mols)} molecules, {batch.npair} pairs, "
        f"{batch.n_slots} slots ({time.perf_counter() - t0:.1f} s)")
    eng = get_engine("cuda")
    eng.path = args.path
    eng.launch_streams = args.launch_streams
    eng.hrr_lds = args.hrr_lds
    dev = DeviceTables(batch)
    schwarz_ms = time_schwarz(eng, dev)
    q = eng.schwarz(dev)
    bound = eng.pair_bounds(dev, q)
    torch.cuda.synchronize()
    plans64 = eng.plan_device(dev, bound, args.threshold, 0.0, stats=True)
    plansmx = eng.plan_device(dev, bound, args.threshold, args.fp32_cutoff, stats=True)
    eris = sum(bk.eris for bk in plans64)
    nquart = sum(bk.n for bk in plans64 if bk.stores())
    nsub = sum(bk.n for bk in plans64)
    n_fp32 = sum(bk.n for bk in plansmx if bk.precision == 1 and bk.stores())
    log(f"plans: {nquart} quartets, {eris} ERIs, {len(plans64)} buckets; "
        f"{n_fp32} quartets FP32-bulk in mixed ({time.perf_counter() - t0:.1f} s)")

    tuner = make_tuner(eng, args)
    dev.out = torch.zeros(batch.n_slots, dtype=torch.float64, device="cuda")
    t64, q64, w64, pb64, last64 = time_steps(eng, dev, bound, args.threshold, 0.0, args.steps,
                                             args.warmup, tuner)
    log(f"fp64: {t64 * 1e3:.1f} ms/step (queue {q64 * 1e3:.1f} ms), {eris / t64:.3e} ERIs/s")
    g64 = None
    if not args.no_graph:
        g64, _ = time_graph(eng, dev, bound, args.threshold, 0.0, args.steps, tuner)
        log(f"fp64 graph replay: {g64 * 1e3:.1f} ms/step, {eris / g64:.3e} ERIs/s")
    out64 = dev.out
    attach_stats(last64, plans64)
    if args.only == "fp64":
        if tuner is not None and args.tune_save:
            tuner.save(args.tune_save)
        roof = roofline_entry(last64, pb64)
        print(json.dumps({"metric": METRIC, "value": eris / t64, "unit": "ERIs/s", "n_gpus": 1,
                          "steps": args.steps, "warmup": args.warmup, "ms_per_step": t64 * 1e3,
                          "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                          "dtype": "f64", "data": "synthetic",
                          "config": {"workload": "FP64 form only", "molecules": args.molecules,
                                     "eris": eris}, "roofline": roof,
                          "graph_ms_per_step": None if g64 is None else g64 * 1e3,
                          "classes_fp64": class_table(last64, pb64)}), flush=True)
        return
    dev.out = torch.zeros(batch.n_slots, dtype=torch.float64, device="cuda")
    tmx, qmx, wmx, pbmx, lastmx = time_steps(eng, dev, bound, args.threshold, args.fp32_cutoff,
                                             args.steps, args.warmup, tuner)
    log(f"mixed: {tmx * 1e3:.1f} ms/step (queue {qmx * 1e3:.1f} ms), {eris / tmx:.3e} ERIs/s")
    gmx = None
    if not args.no_graph:
        gmx, _ = time_graph(eng, dev, bound, args.threshold, args.fp32_cutoff, args.steps, tuner)
        log(f"mixed graph replay: {gmx * 1e3:.1f} ms/step, {eris / gmx:.3e} ERIs/s")
    outmx = dev.out
    # bookkeeping of the buckets actually launched (after the autotuner's FP32 mask)
    plansmx = eng.plan_device(dev, bound, args.threshold, args.fp32_cutoff, stats=True)
    attach_stats(lastmx, plansmx)
    diff = outmx.sub_(out64).abs_()
    max_err = float(diff.max())
    rms_err = float(torch.linalg.vector_norm(diff)) / np.sqrt(diff.numel())
    del diff, outmx
    roof = roofline_entry(last64, pb64)
    roof_mx = roofline_entry(lastmx, pbmx)
    # the headline step is the HIP-graph replay of the whole step (queue fill and every
    # launch) when captured, else the eager step; both are reported
    tmx_head = gmx if gmx is not None else tmx
    t64_head = g64 if g64 is not None else t64
    result = {
        "metric": METRIC, "value": eris / tmx_head, "unit": "ERIs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": tmx_head * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64+f32",
        "dtype_detail": mixed_detail(lastmx, t64_head, tmx_head),
        "data": "synthetic",
        "config": {"workload": f"batch of {args.molecules} molecules, 4-24 atoms, s/p/d "
                               "Cartesian, K 1-6, exps [0.05,1e3], packed 8-fold output",
                   "molecules": args.molecules, "shell_quartets": nquart, "eris": eris,
                   "schwarz_threshold": args.threshold, "fp32_cutoff": args.fp32_cutoff,
                   "fp32_quartets": n_fp32, "kernel_path": args.path,
                   "autotuned": not args.no_tune,
                   "fp32_quartets_run_fp32": sum(bk.n for bk in lastmx
                                                 if bk.launch_precision() == 1
                                                 and bk.stores()),
                   "sub_quartets": nsub},
        "step_timed_as": "hip graph replay" if gmx is not None else "eager launches",
        "fp64": {"value": eris / t64_head, "ms_per_step": t64_head * 1e3},
        "eager_ms_per_step": {"fp64": t64 * 1e3, "mixed": tmx * 1e3},
        "graph_ms_per_step": {"fp64": None if g64 is None else g64 * 1e3,
                              "mixed": None if gmx is None else gmx * 1e3},
        "queue_build_ms": {"fp64": q64 * 1e3, "mixed": qmx * 1e3},
        "schwarz_ms": schwarz_ms,
        "fp32_quartet_fraction": fp32_fraction(lastmx),
        "mixed_vs_fp64": {"max_abs_err": max_err, "rms_err": rms_err},
        "wall_ms_per_step": {"fp64": w64 * 1e3, "mixed": wmx * 1e3},
        "roofline": roof,
        "roofline_mixed": roof_mx,
        "classes_fp64": class_table(last64, pb64),
        "classes_mixed": class_table(lastmx, pbmx),
    }
    if not args.no_cpu:
        result["cpu_baseline"], result["cpu_baseline_one_core"] = cpu_baselines(batch, plans64)
    del dev, out64
    torch.cuda.empty_cache()
    if not args.no_workloads:
        # the single-molecule workloads, one warm-up and a few timed steps each
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.no_cpu = args.workload_steps, 1, True
        result["workloads"] = {}
        for name in CONFIG_WORKLOADS:
            t0 = time.perf_counter()
            result["workloads"][name] = workload_summary(run_workload(name, sub, tuner))
            torch.cuda.empty_cache()
            log(f"workload {name}: {time.perf_counter() - t0:.0f} s")
    if tuner is not None and args.tune_save:
        tuner.save(args.tune_save)
    if not args.no_serving:
        result["serving"] = serving_bench(args.serving_requests, args.serving_saturation,
                                          args.threshold, args.fp32_cutoff,
                                          args.serving_streams,
                                          max_batch=args.serving_max_batch,
                                          batch_flops=args.serving_batch_flops or None)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
