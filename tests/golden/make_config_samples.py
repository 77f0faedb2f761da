"""Write tests/golden/config_samples.npz: oracle blocks of sampled shell quartets of
the dense s/s/p/d (N=44), large s/p/d/f (N=320) and high-l pure s-g (N=150)
workloads.

For every canonical angular-momentum class present among the unique shell
quartets (i >= j, k >= l, ij >= kl) of a workload, up to ``per_class`` quartets
are drawn (seeded), and their final normalised blocks (norm_cont, component
order, Cartesian->pure as the shells ask) are evaluated by the oracle four
times: in IEEE binary128 with every primitive pair built on the centre of its
larger exponent ("quad", oracle/eri_hp.c: the accuracy reference), in x87
extended precision with the bounded-coefficient ket build ("exact"), with the
reference algorithm in FP64 ("fp64"), and with the bounded-coefficient ket
build in plain FP64 ("kvrr64").  The GPU tests evaluate the same quartets and
compare.  Run once in a container with the oracle's dependencies; the x87 and
FP64 columns take about twenty minutes on 8 cores, the binary128 column about
as long.  ``--quad-only`` keeps the other columns of an existing file and
(re)computes the binary128 one.
"""

import os
import sys
from multiprocessing import Pool

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from gbasis_amd import workloads  # noqa: E402
import oracle.assembly as assembly  # noqa: E402
from oracle.assembly import normalized_block  # noqa: E402
from oracle.os_hgp import os_hgp_quartet  # noqa: E402

CONFIGS = {
    # name: (basis builder, samples per class)
    "dense_small": (workloads.dense_small, 3),
    "large_spdf": (workloads.large_spdf, 2),
    "high_l_pure": (workloads.high_l_pure, 1),
}


def _canon(ls):
    la, lb, lc, ld = ls
    a, b = max(la, lb), min(la, lb)
    c, d = max(lc, ld), min(lc, ld)
    return max((a, b, c, d), (c, d, a, b))


def pick(basis, per_class, seed):
    n = len(basis)
    pairs = [(i, j) for i in range(n) for j in range(i + 1)]
    by_cls = {}
    for p, (i, j) in enumerate(pairs):
        for k, l in pairs[:p + 1]:
            key = _canon([basis[x].angmom for x in (i, j, k, l)])
            by_cls.setdefault(key, []).append((i, j, k, l))
    rng = np.random.default_rng(seed)
    out = []
    for key in sorted(by_cls):
        lst = by_cls[key]
        for t in rng.choice(len(lst), size=min(per_class, len(lst)), replace=False):
            out.append(lst[int(t)])
    return out


def _eval(args):
    name, q, quad_only = args
    basis = CONFIGS[name][0]()
    cts = [s.coord_type for s in basis]
    qd = normalized_block(basis, cts, *q, exact="quad")
    if quad_only:
        return name, q, qd, None, None, None
    ex = normalized_block(basis, cts, *q, exact=True)
    f64 = normalized_block(basis, cts, *q, exact=False)
    k64 = _kvrr64_block(basis, cts, q)
    return name, q, qd, ex, f64, k64


def _kvrr64_engine(s1, s2, s3, s4, exact=False):
    return os_hgp_quartet(*[x for s in (s1, s2, s3, s4)
                            for x in (s.coord, s.angmom, s.angmom_components_cart, s.exps,
                                      s.coeffs)], ket="vrr")


def _kvrr64_block(basis, cts, q):
    """normalized_block with the oracle engine swapped for the FP64 two-centre build."""
    saved = assembly._engine
    assembly._engine = _kvrr64_engine
    try:
        return normalized_block(basis, cts, *q, exact=False)
    finally:
        assembly._engine = saved


def main():
    quad_only = "--quad-only" in sys.argv
    path = os.path.join(HERE, "config_samples.npz")
    arrays = {}
    if quad_only:
        with np.load(path, allow_pickle=False) as z:
            arrays = {k: z[k] for k in z.files}
    jobs = []
    for seed, (name, (builder, per)) in enumerate(CONFIGS.items()):
        picked = pick(builder(), per, 100 + seed)
        if quad_only:
            assert [tuple(q) for q in arrays[f"{name}_idx"].tolist()] == sorted(picked), name
        for q in picked:
            jobs.append((name, q, quad_only))
    # heaviest first
    jobs.sort(key=lambda j: -sum(CONFIGS[j[0]][0]()[x].angmom for x in j[1]))
    res = {}
    with Pool(8) as pool:
        for out in pool.imap_unordered(_eval, jobs):
            name, q = out[0], out[1]
            res.setdefault(name, []).append((q,) + out[2:])
            print(name, q, out[2].shape, flush=True)
    for name, items in res.items():
        items.sort(key=lambda t: t[0])
        arrays[f"{name}_idx"] = np.array([t[0] for t in items], dtype=np.int32)
        arrays[f"{name}_shape"] = np.array([t[1].shape for t in items], dtype=np.int32)
        arrays[f"{name}_quad"] = np.concatenate([t[1].ravel() for t in items])
        if not quad_only:
            arrays[f"{name}_exact"] = np.concatenate([t[2].ravel() for t in items])
            arrays[f"{name}_fp64"] = np.concatenate([t[3].ravel() for t in items])
            arrays[f"{name}_kvrr64"] = np.concatenate([t[4].ravel() for t in items])
    np.savez_compressed(path, **arrays)


if __name__ == "__main__":
    main()
