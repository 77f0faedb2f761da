"""Device residency and kernel dispatch for the ERI path.

``DeviceTables`` holds one batch of molecules' tables in HBM (torch tensors are
only the allocation mechanism; every computation runs in the HIP kernels of
``libgbasis_amd.so``).  ``EriEngine`` runs the three device stages:

1. Schwarz bounds Q_ij = sqrt(max |(ij|ij)|) for every shell pair
   (gbasis/integrals/_schwarz_screening.py:14-90), one launch per class;
2. the screened, class-bucketed work queue, built on the device
   (``plan_device``; ``plan_quartets`` is the host enumeration kept for checks);
3. one launch per (class, primitive-count bin, precision) writing the
   canonical packed array.

With ``device_counts`` the launches of stage 3 read their quartet counts from
the queue builder's cursors on the device, so a whole step is enqueued without
a host round trip and can be recorded once into a HIP graph (``capture``).
"""

import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from ._lib import (GBX_FP32_BULK, GBX_FP64, GBX_MODE_PACKED, GBX_MODE_PACKED_ADD,
                   GBX_MODE_SCHWARZ, GBX_POT_COULOMB, GBX_POT_ERF, GBX_POT_ERFC, GbxProgram,
                   GbxTables, check, ptr)
from .boys import BOYS_H, boys_table
from .programs import canonical_class, encoded_program
from .roofline import class_flops, quartet_flops
from .tables import (NUM_BUCKETS, STORE_PHASES, BatchTables, MoleculeTables, decode_head,
                     f2norm_table)

DEFAULT_SCRATCH_BYTES = 2 << 30


@dataclass
class Bucket:
    """One launch: quartets of one class / precision / primitive-count bin."""
    cls: tuple
    nprimq: int          # largest primitive-quartet count in the bucket
    precision: int       # GBX_FP64 or GBX_FP32_BULK
    dense: bool          # needs a dense Cartesian->pure output transform
    quart: object        # device int32 (n, 2)
    n: int               # shell quartets
    eris: int            # canonical packed slots written
    flops: float         # algorithmic FLOPs (roofline.quartet_flops)
    path: str = ""       # "reg", "split" or "coop" once launched
    run_precision: int = -1   # precision actually launched (autotuner); -1: = precision
    variant: tuple = None     # autotuned variant: ("R", lanes per quartet, register HRR)
                              # or ("coop",) for the cooperative kernel
    seg: tuple = None         # fused classes: (device segment offsets, device counts, max lanes)
    count: object = None      # device int32 count (n is then the capacity); None: n exact
    key: tuple = None         # ("head", class/dense/precision head) for a fused launch,
                              # ("key", bucket key) otherwise
    phase: int = 0            # sub-quartet phase: 0 stores the packed slots, 1-3 add to
                              # them and run after the phases before

    def mode(self):
        """Packed output mode of this bucket's launch."""
        return GBX_MODE_PACKED if self.stores() else GBX_MODE_PACKED_ADD

    def stores(self):
        """True for the first sub-quartet of each of its shell quartets (phase 0): the
        launch that stores, and the one that counts the shell quartet."""
        return self.phase in STORE_PHASES

    def launch_precision(self):
        return self.precision if self.run_precision < 0 else self.run_precision

    def lanes(self):
        return self.variant[1] if self.variant and self.variant[0] == "R" else None

    def hrr(self):
        """Autotuned finishing pass of a split class: True register HRR kernel, False the
        cooperative pass, None the engine default."""
        if self.variant and self.variant[0] == "R" and len(self.variant) > 2:
            return bool(self.variant[2])
        return None

    def host_quartets(self):
        """(n, 2) int32 NumPy copy of this bucket's quartets (gathers the segments of a
        segmented bucket)."""
        if self.seg is None:
            q = self.quart.cpu().numpy()
            return q if self.count is None else q[:int(self.count.item())]
        offs = self.seg[0].cpu().numpy()
        cnts = self.seg[1].cpu().numpy()
        q = self.quart.cpu().numpy()
        parts = [q[o:o + c] for o, c in zip(offs.tolist(), cnts.tolist()) if c > 0]
        return np.concatenate(parts) if parts else np.zeros((0, 2), np.int32)


def _require_gpu():
    if not torch.cuda.is_available():
        raise RuntimeError("gbasis_amd: no HIP device visible; the ERI path runs only on the GPU")


class ProgramCache:
    """Class programs resident on the device, built once per class."""

    def __init__(self, device):
        self.device = device
        self._cache = {}

    def get(self, cls, sub=None):
        """(device struct, device arrays, header) of a class, or with
        ``sub = (parts, k)`` of the k-th of its ``parts`` primitive-pass programs."""
        key = cls if sub is None else (cls, sub)
        if key not in self._cache:
            header, items, phases, out_slots = encoded_program(cls, sub)
            d_items = torch.from_numpy(np.ascontiguousarray(items) if items.size else
                                       np.zeros(4, np.int32)).to(self.device)
            d_phases = torch.from_numpy(np.ascontiguousarray(phases)).to(self.device)
            d_out = torch.from_numpy(np.ascontiguousarray(out_slots)).to(self.device)
            st = GbxProgram()
            st.items = ptr(d_items)
            st.phases = ptr(d_phases)
            st.out_slots = ptr(d_out)
            for i in range(16):
                st.header[i] = int(header[i])
            self._cache[key] = (st, (d_items, d_phases, d_out), header)
        return self._cache[key]


def _h2d(a, device):
    """Asynchronous host-to-device copy through pinned memory (stream-ordered on the
    current stream; torch keeps the pinned block until the copy is done)."""
    src = torch.from_numpy(np.ascontiguousarray(a))
    return src.pin_memory().to(device, non_blocking=True)


class DeviceTables:
    """A BatchTables object copied to the device, plus its gbx_tables struct.

    All tables travel in one pinned blob and one asynchronous copy on the
    current stream.  ``mol_fp32_cutoff`` (optional, one value per molecule)
    selects mixed precision per molecule.
    """

    ALIGN = 256

    POTENTIALS = {"coulomb": GBX_POT_COULOMB, "erf": GBX_POT_ERF, "erfc": GBX_POT_ERFC}

    def __init__(self, batch: BatchTables, device="cuda", mol_fp32_cutoff=None,
                 potential="coulomb", omega=None):
        _require_gpu()
        if potential not in self.POTENTIALS:
            raise ValueError(f"potential must be one of {tuple(self.POTENTIALS)}")
        if potential != "coulomb" and (omega is None or not omega > 0):
            raise ValueError("omega > 0 is required for the erf and erfc potentials")
        if potential != "coulomb" and mol_fp32_cutoff is not None:
            raise ValueError("range-separated potentials are evaluated in FP64 only")
        self.potential = potential
        self.batch = batch
        self.device = torch.device(device)
        # largest output dimension of the batch's shells per angular momentum
        info = np.asarray(batch.shell_info).reshape(-1, 8)
        self.max_out_dim = {int(l): int(info[info[:, 1] == l, 5].max())
                            for l in np.unique(info[:, 1])}
        tab = boys_table()
        arrays = {
            "shell_info": batch.shell_info.astype(np.int32),
            "comp_map": batch.comp_map.astype(np.int32),
            "comp_scale": batch.comp_scale.astype(np.float64),
            "out_mat": batch.out_mat.astype(np.float64),
            "pair_info": batch.pair_info.astype(np.int32),
            "pair_geom": batch.pair_geom.astype(np.float64),
            "pair_out_base": batch.pair_out_base.view(np.int64),
            "prim_pairs": batch.prim_pairs.astype(np.float64),
            "f2norm": f2norm_table(8),
            "pair_key": batch.pair_key.astype(np.int32),
            "pair_sub": batch.pair_sub.astype(np.int32),
            "pair_npf": batch.pair_npf.astype(np.int32),
            "mol_cand_off": batch.mol_cand_off.astype(np.int64),
            "mol_pair_off": np.asarray(batch.mol_pair_off, dtype=np.int32),
            "boys": tab,
        }
        if mol_fp32_cutoff is not None:
            arrays["mol_fp32_cutoff"] = np.asarray(mol_fp32_cutoff, dtype=np.float64).reshape(
                len(batch.mols))
        # bra == ket pair lists of the Schwarz pass, one per (class, primitive bin)
        self.schwarz_lists = []
        for i, (cls, nprimq, _, quart) in enumerate(batch.schwarz_buckets()):
            arrays[f"sw{i}"] = quart
            self.schwarz_lists.append((cls, nprimq, f"sw{i}"))
        offs, total = {}, 0
        for k, v in arrays.items():
            offs[k] = total
            total += -(-max(v.nbytes, 1) // self.ALIGN) * self.ALIGN
        host = torch.empty(total, dtype=torch.uint8, pin_memory=True)
        hv = host.numpy()
        for k, v in arrays.items():
            hv[offs[k]:offs[k] + v.nbytes] = np.ascontiguousarray(v).reshape(-1).view(np.uint8)
        self.blob = host.to(self.device, non_blocking=True)
        dt = {np.dtype(np.int32): torch.int32, np.dtype(np.int64): torch.int64,
              np.dtype(np.float64): torch.float64}
        for k, v in arrays.items():
            view = self.blob[offs[k]:offs[k] + v.nbytes].view(dt[v.dtype]).view(v.shape)
            setattr(self, k, view)
        self.schwarz_lists = [(c, n, getattr(self, k)) for c, n, k in self.schwarz_lists]
        if mol_fp32_cutoff is None:
            self.mol_fp32_cutoff = None
        st = GbxTables()
        st.shell_info = ptr(self.shell_info)
        st.comp_map = ptr(self.comp_map)
        st.comp_scale = ptr(self.comp_scale)
        st.out_mat = ptr(self.out_mat)
        st.pair_info = ptr(self.pair_info)
        st.pair_geom = ptr(self.pair_geom)
        st.pair_out_base = ptr(self.pair_out_base)
        st.prim_pairs = ptr(self.prim_pairs)
        st.f2norm = ptr(self.f2norm)
        st.boys_table = ptr(self.boys)
        st.boys_nt = tab.shape[0]
        st.boys_nord = tab.shape[1]
        st.boys_h = BOYS_H
        st.potential = self.POTENTIALS[potential]
        st.omega = float(omega) if omega is not None else 0.0
        self.struct = st
        # Schwarz bounds are Coulomb whatever the operator (as the reference's screener)
        if potential == "coulomb":
            self.struct_coulomb = st
        else:
            sc = GbxTables()
            ctypes.pointer(sc)[0] = st
            sc.potential = GBX_POT_COULOMB
            sc.omega = 0.0
            self.struct_coulomb = sc
        self.queue = {}   # work-queue buffers of plan_device, per layout


class EriEngine:
    """Runs the ERI kernels for one device; holds programs and scratch."""

    def __init__(self, device="cuda", scratch_bytes=DEFAULT_SCRATCH_BYTES):
        _require_gpu()
        self.lib = _lib.load()
        self.device = torch.device(device)
        self.programs = ProgramCache(self.device)
        self.scratch_bytes = int(scratch_bytes)
        self._scratch = {}
        self._ws = {}
        self._streams = {}
        self.launch_streams = 1
        # default finishing pass of the split classes (the autotuner picks per bucket):
        # False = the cooperative HRR pass, True = the register HRR kernels
        self.hrr_reg = False
        # stream_workspaces: split accumulators and work-queue buffers come from
        # grow-only buffers per stream instead of fresh allocations (a server running
        # batches on several streams then allocates nothing per batch)
        self.stream_workspaces = False
        self._reg = {}
        self._hrr = {}
        self._split = {}
        self._pf = {}
        self._rys_dev = None
        self._rys_idx = {}
        # int8 device mask per class*2+dense: classes allowed the FP32 recursion bulk.
        # By default only classes with every l <= 2: their FP32 error against FP64 is
        # measured within the 1e-8 bound at the default cutoff; f and g shells with wide
        # exponent ranges are not (errors near 1e-7 seen), and the autotuner only ever
        # narrows this set.
        self.fp32_ok = None
        # every mask the engine has used, by version: a work queue sized for one mask
        # (its cached layout carries the version) and a HIP graph captured with it keep
        # reading that mask's device copy, so none of them is ever freed
        self._masks = []
        self.set_fp32_allowed({})
        self.path = "auto"   # "coop" forces the cooperative kernel for every class
        # horizontal-recursion pass in LDS (gbx_eri_hrr_lds): "auto" after every Rys pass
        # (whose accumulator writes are contiguous in the quartet-major layout only that
        # pass reads), after the split kernels where its 256- or 512-lane forms run, and
        # for the other classes whose cooperative pass would run on global scratch;
        # "always" wherever it fits, "never"
        self.hrr_lds = "auto"
        self._recipes = {}
        # deterministic: variant choices depend only on (class, primitive bin, precision),
        # never on how many quartets a launch holds, so the value of every packed slot
        # is independent of the batch it runs in (serving)
        self.deterministic = False

    def scratch(self):
        """Global scratch of the cooperative kernel, one buffer per stream."""
        key = torch.cuda.current_stream(self.device).cuda_stream
        if key not in self._scratch:
            self._scratch[key] = torch.empty(self.scratch_bytes, dtype=torch.uint8,
                                             device=self.device)
        return self._scratch[key]

    def workspace(self, name, numel, dtype):
        """Grow-only device buffer ``name`` of the current stream, at least ``numel``
        elements of ``dtype``; its contents are only valid until the next use."""
        key = (torch.cuda.current_stream(self.device).cuda_stream, name)
        buf = self._ws.get(key)
        if buf is None or buf.numel() < numel or buf.dtype != dtype:
            size = max(int(numel), int(1.25 * (buf.numel() if buf is not None else 0)), 1)
            buf = torch.empty(size, dtype=dtype, device=self.device)
            self._ws[key] = buf
        return buf[:numel]

    def _stream(self):
        return ctypes.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    LDS_TARGET = 32768   # bytes per workgroup: ~5 single-wave workgroups per CU

    def _workspace(self, cls, precision, pf, prim_only, sub=None):
        st, _, _ = self.programs.get(cls, sub)
        lds = ctypes.c_int64()
        glb = ctypes.c_int64()
        check(self.lib.gbx_coop_workspace(ctypes.byref(st), *cls, precision, pf, int(prim_only),
                                          ctypes.byref(lds), ctypes.byref(glb)),
              "gbx_coop_workspace")
        return lds.value, glb.value

    def _plan(self, cls, nprimq, precision, prim_only=False, sub=None):
        """Primitive quartets in flight (PF) for the cooperative kernel.

        Largest power of two <= min(nprimq, 64) whose LDS workspace stays within
        LDS_TARGET; if even PF = 1 does not fit the LDS of one workgroup (160 KiB on
        gfx950; workspaces above 40 KiB run 4-16 wavefronts) the kernel runs on global
        scratch, where PF = min(nprimq, 16) keeps the wave busy (PF = 1 for the
        split primitive pass, whose parallelism comes from the parts).
        """
        key = (cls, nprimq, precision, prim_only, sub)
        if key in self._pf:
            return self._pf[key]
        pf = 1
        while pf < min(nprimq, 64):
            pf *= 2
        while pf > 1 and self._workspace(cls, precision, pf, prim_only, sub)[0] > self.LDS_TARGET:
            pf //= 2
        if self._workspace(cls, precision, pf, prim_only, sub)[1] > 0:
            pf = 1
            while not prim_only and pf < min(nprimq, 16):
                pf *= 2
        self._pf[key] = pf
        return pf

    SUB_PARTS = (2, 3, 4, 6, 8)

    def _prim_parts(self, cls, precision):
        """Number of primitive-pass programs of a class: 1 if its primitive store fits
        one workgroup's LDS, else the fewest accumulator subsets whose stores all fit
        (1 again, on global scratch, if none does)."""
        key = ("parts", cls, precision)
        if key not in self._pf:
            parts = 1
            if self._workspace(cls, precision, 1, True)[1] > 0:
                for S in self.SUB_PARTS:
                    if all(self._workspace(cls, precision, 1, True, (S, k))[1] == 0
                           for k in range(S)):
                        parts = S
                        break
            self._pf[key] = parts
        return self._pf[key]

    COOP_TARGET_BLOCKS = 8192   # workgroups that keep every CU busy on the primitive pass
    COOP_MIN_CHUNK = 16         # primitive quartets per part at least

    def _coop_parts(self, nq, nprimq):
        """Parts per shell quartet of the cooperative primitive pass: enough workgroups to
        fill the GPU, at least COOP_MIN_CHUNK primitive quartets each.  In deterministic
        mode the count depends on the primitive bin only."""
        cap = max(1, nprimq // self.COOP_MIN_CHUNK)
        R = 1
        if self.deterministic:
            while R * 2 <= min(cap, 16):
                R *= 2
            return R
        while R * 2 <= min(cap, 256) and nq * R < self.COOP_TARGET_BLOCKS:
            R *= 2
        return R

    TARGET_LANES = 1 << 18   # lanes in flight that keep all 1024 SIMDs of the GPU busy

    DETERMINISTIC_LANES = 8

    def _lanes_per_quartet(self, nq, nprimq):
        """R lanes per shell quartet for the split kernels: enough lanes to fill the GPU,
        never more than the (upper bound of the) primitive quartets per shell quartet.
        In deterministic mode R depends on the primitive bin only."""
        R = 1
        if self.deterministic:
            while R < min(nprimq, self.DETERMINISTIC_LANES):
                R *= 2
            return R
        while R * 2 <= nprimq and nq * R * 2 <= self.TARGET_LANES:
            R *= 2
        return R

    def split_count(self, cls):
        if cls not in self._split:
            self._split[cls] = int(self.lib.gbx_split_count(*cls))
        return self._split[cls]

    def has_hrr(self, cls):
        if cls not in self._hrr:
            self._hrr[cls] = bool(self.lib.gbx_has_hrr_kernel(*cls))
        return self._hrr[cls]

    def has_reg(self, cls):
        if cls not in self._reg:
            self._reg[cls] = bool(self.lib.gbx_has_reg_kernel(*cls))
        return self._reg[cls]

    def launch_bucket(self, dev, bk, out, mode=None, precision=None, lanes=None,
                      hrr=None, path=None):
        """Launch one bucket; ``mode`` defaults to the bucket's own (store in phase 0,
        add in the later phases)."""
        prec = bk.launch_precision() if precision is None else precision
        rys_layout = rys_batch = None
        if path is None:
            path = (bk.variant[0] if bk.variant and bk.variant[0] in ("coop", "rys")
                    else self.path)
            if bk.variant and bk.variant[0] == "rys":
                rys_layout = bk.variant[1] if len(bk.variant) > 1 else None
                rys_batch = bk.variant[2] if len(bk.variant) > 2 else None
        mode = bk.mode() if mode is None else mode
        return self.launch_class(dev, bk.cls, bk.quart, bk.nprimq, prec, mode, out, bk.dense,
                                 path, lanes if lanes is not None else bk.lanes(), bk.seg,
                                 bk.count, hrr if hrr is not None else bk.hrr(), rys_layout,
                                 rys_batch)

    def launch_class(self, dev: DeviceTables, cls, quartets, nprimq, precision, mode, out,
                     dense=False, path="auto", lanes=None, seg=None, count=None, hrr=None,
                     rys_layout=None, rys_batch=None):
        """One launch for a (class, bin) bucket; quartets is a device int32 (n,2).

        Classes with a generated register kernel and no dense output transform run
        one quartet per lane (split register kernels plus a horizontal-recursion pass
        for the classes too large for one lane); all others run the cooperative
        kernel.  ``count``: device int32 holding the number of quartets (n is then
        the capacity).
        """
        packed = mode in (GBX_MODE_PACKED, GBX_MODE_PACKED_ADD)
        tab = dev.struct if packed else dev.struct_coulomb
        if dev.potential != "coulomb" and packed:
            # range-separated operators: cooperative kernels, FP64
            path = "coop"
            if precision != GBX_FP64:
                raise ValueError("range-separated potentials are evaluated in FP64 only")
        if seg is not None and (path == "coop" or dense or not self.has_reg(cls)):
            # a fused launch covers several primitive bins of the queue: only the register
            # kernels read its segment table
            raise ValueError(f"fused launch of class {cls} planned for a register kernel "
                             "cannot run on the cooperative kernel; plan again")
        if path not in ("coop", "rys") and not dense and self.has_reg(cls):
            if seg is not None:
                check(self.lib.gbx_eri_reg_segments(ctypes.byref(tab), *cls, ptr(quartets),
                                                    ptr(seg[0]), ptr(seg[1]), seg[2], precision,
                                                    mode, ptr(out), self._stream()),
                      f"gbx_eri_reg_segments{cls}")
            else:
                check(self.lib.gbx_eri_reg(ctypes.byref(tab), *cls, ptr(quartets),
                                           quartets.shape[0], ptr(count), precision, mode,
                                           ptr(out), self._stream()), f"gbx_eri_reg{cls}")
            return "reg"
        if quartets.shape[0] == 0:
            # a bucket of capacity 0 (a class and bin with no quartet in this precision)
            return "empty"
        if path == "auto" and self.rys_default(cls) and dev.potential == "coulomb":
            path = "rys"
        if path == "rys":
            if not (self.rys_ok(cls) and dev.potential == "coulomb"):
                raise ValueError(f"class {cls} has no Rys primitive pass")
            return self._launch_rys(dev, tab, cls, quartets, nprimq, mode, out, count,
                                    self.rys_layout(cls) if rys_layout is None else rys_layout,
                                    rys_batch)
        st, _, _ = self.programs.get(cls)
        nsplit = self.split_count(cls)
        if path != "coop" and nsplit > 0:
            nq = quartets.shape[0]
            nacc = st.header[1]
            R = lanes if lanes else self._lanes_per_quartet(nq, nprimq)
            if self.stream_workspaces:
                acc = self.workspace("acc", nacc * R * nq, torch.float64)
            else:
                acc = torch.empty(nacc * R * nq, dtype=torch.float64, device=self.device)
            for k in range(nsplit):
                check(self.lib.gbx_eri_split(ctypes.byref(tab), *cls, k, ptr(quartets), nq,
                                             ptr(count), precision, R, ptr(acc), nq,
                                             self._stream()), f"gbx_eri_split{cls}[{k}]")
            use_reg = self.hrr_reg if hrr is None else hrr
            if packed and not dense and use_reg and self.has_hrr(cls):
                check(self.lib.gbx_eri_hrr_reg(ctypes.byref(tab), *cls, ptr(quartets), nq,
                                               ptr(count), ptr(acc), nq, R, mode, ptr(out),
                                               self._stream()), f"gbx_eri_hrr_reg{cls}")
                return "split"
            self._hrr_pass(dev, tab, st, cls, quartets, count, mode, acc, R, out,
                           lds=self._lds_hrr_plan(dev, cls, mode, split=True))
            return "split"
        nq = quartets.shape[0]
        R = self._coop_parts(nq, nprimq)
        full_fits = self._workspace(cls, precision, 1, False)[1] == 0
        sc = self.scratch()
        if R == 1 and full_fits:
            pf = self._plan(cls, nprimq, precision)
            check(self.lib.gbx_eri_coop(ctypes.byref(tab), ctypes.byref(st), *cls,
                                        ptr(quartets), nq, ptr(count), precision, pf, mode,
                                        ptr(out), ptr(sc), sc.numel(), self._stream()),
                  f"gbx_eri_coop{cls}")
            return "coop"
        # two passes: the primitive work split over R workgroups per quartet (primitive
        # workspace only), then the horizontal recursion and epilogue.  A class whose
        # primitive store exceeds the LDS runs its primitive pass as several programs
        # over subsets of the accumulators, each of which fits.
        nacc = st.header[1]
        if self.stream_workspaces:
            acc = self.workspace("acc", nacc * R * nq, torch.float64)
        else:
            acc = torch.empty(nacc * R * nq, dtype=torch.float64, device=self.device)
        parts = self._prim_parts(cls, precision)
        subs = [None] if parts == 1 else [(parts, k) for k in range(parts)]
        for sub in subs:
            stp = st if sub is None else self.programs.get(cls, sub)[0]
            pf = self._plan(cls, max(1, -(-nprimq // R)), precision, prim_only=True, sub=sub)
            check(self.lib.gbx_eri_coop_acc(ctypes.byref(tab), ctypes.byref(stp), *cls,
                                            ptr(quartets), nq, ptr(count), precision, pf, R,
                                            ptr(acc), nq, ptr(sc), sc.numel(), self._stream()),
                  f"gbx_eri_coop_acc{cls}")
        self._hrr_pass(dev, tab, st, cls, quartets, count, mode, acc, R, out)
        return "coop2"

    # Rys primitive pass (gbx_eri_rys_acc): admissible for 6 <= L <= 17 (the autotuner
    # times it against the other variants), taken by default from RYS_MIN_L on for the
    # classes without register kernels
    RYS_MIN_L = 8

    def rys_ok(self, cls):
        la, lb, lc, ld = cls
        return (6 <= la + lb + lc + ld <= 17 and lb <= 4 and ld <= 4 and la + lb <= 8
                and lc + ld <= 8 and (la + lb + 1) * (lc + ld + 1) <= 128)

    RYS_STRIDED, RYS_GROUPED = 0, 1

    def rys_layout(self, cls):
        """Default accumulator layout of the Rys pass: groups sharing Ix Iy where they are
        small ((l_b + 1)(l_d + 1) < 15), one accumulator per lane slot otherwise (measured
        faster on the g classes; the autotuner times both)."""
        return self.RYS_GROUPED if (cls[1] + 1) * (cls[3] + 1) < 15 else self.RYS_STRIDED

    def _rys_strided(self, cls):
        """Per accumulator of the class program (order (ket c, bra a)): the offsets of
        (a_i, c_i) in the (l_a + l_b + 1) x (l_c + l_d + 1) tables, x | y << 10 | z << 20."""
        key = ("strided", cls)
        if key not in self._rys_idx:
            from .programs import comps

            la, lb, lc, ld = cls
            W = lc + ld + 1
            vals = [(a[0] * W + c[0]) | ((a[1] * W + c[1]) << 10) | ((a[2] * W + c[2]) << 20)
                    for nc in range(lc, lc + ld + 1) for c in comps(nc)
                    for na in range(la, la + lb + 1) for a in comps(na)]
            self._rys_idx[key] = (torch.tensor(vals, dtype=torch.int32, device=self.device),
                                  len(vals))
        return self._rys_idx[key]

    def rys_default(self, cls):
        return self.rys_ok(cls) and sum(cls) >= self.RYS_MIN_L and self.split_count(cls) == 0 \
            and not self.has_reg(cls)

    def _rys_table(self):
        if self._rys_dev is None:
            from .rys import rys_table

            tab = rys_table()
            if tab.size != int(self.lib.gbx_rys_table_size()):
                raise RuntimeError("Rys node table does not match the library's layout")
            self._rys_dev = torch.from_numpy(tab).to(self.device)
        return self._rys_dev

    def _rys_groups(self, cls):
        """Lane slots of the Rys pass for a class (gbx_eri_rys_acc): the accumulators
        [a0|c0] grouped by (a_x, a_y, c_x, c_y) -- a group's members differ in a_z and c_z
        only and share the product Ix Iy -- each group a lane, groups of one shape
        (numbers of a_z and c_z) filling whole wavefronts.  Returns (device int32
        [descriptors, member rows], number of slots)."""
        if cls not in self._rys_idx:
            from .programs import comps

            la, lb, lc, ld = cls
            Lab, Lcd = la + lb, lc + ld
            W, MA, MC = Lcd + 1, lb + 1, ld + 1
            fidx = {}
            for nc in range(lc, Lcd + 1):
                for c in comps(nc):
                    for na in range(la, Lab + 1):
                        for a in comps(na):
                            fidx[(a, c)] = len(fidx)
            by_shape = {}
            for ax in range(Lab + 1):
                for ay in range(Lab + 1 - ax):
                    az0, az1 = max(0, la - ax - ay), Lab - ax - ay
                    for cx in range(Lcd + 1):
                        for cy in range(Lcd + 1 - cx):
                            cz0, cz1 = max(0, lc - cx - cy), Lcd - cx - cy
                            if az1 < az0 or cz1 < cz0:
                                continue
                            na_, nc_ = az1 - az0 + 1, cz1 - cz0 + 1
                            desc = ((ax * W + cx) | ((ay * W + cy) << 7) | (az0 << 14)
                                    | (na_ << 18) | (cz0 << 22) | (nc_ << 26))
                            rows = [-1] * (MA * MC)
                            for i in range(na_):
                                for j in range(nc_):
                                    rows[i * MC + j] = fidx[((ax, ay, az0 + i), (cx, cy, cz0 + j))]
                            by_shape.setdefault((na_, nc_), []).append((desc, rows))
            descs, rows = [], []
            for shape in sorted(by_shape, reverse=True):
                grp = by_shape[shape]
                pad = -len(grp) % 64
                grp = grp + [(-1, [-1] * (MA * MC))] * pad
                descs += [d for d, _ in grp]
                rows += [r for _, r in grp]
            pad = -len(descs) % 512
            descs += [-1] * pad
            rows += [[-1] * (MA * MC)] * pad
            members = sum(1 for r in rows for f in r if f >= 0)
            assert members == len(fidx), (cls, members, len(fidx))
            arr = np.concatenate([np.array(descs, np.int64),
                                  np.array(rows, np.int64).reshape(-1)]).astype(np.int32)
            self._rys_idx[cls] = (torch.from_numpy(arr).to(self.device), len(descs))
        return self._rys_idx[cls]

    def _launch_rys(self, dev, tab, cls, quartets, nprimq, mode, out, count, layout,
                    batch=None):
        """Rys primitive pass split over R workgroups per quartet, then the cooperative
        horizontal recursion and epilogue (FP64).  ``batch``: primitive quartets per LDS
        batch, at most (None: the library's default)."""
        nq = quartets.shape[0]
        R = self._coop_parts(nq, nprimq)
        st, _, _ = self.programs.get(cls)
        nacc = st.header[1]
        if self.stream_workspaces:
            acc = self.workspace("acc", nacc * R * nq, torch.float64)
        else:
            acc = torch.empty(nacc * R * nq, dtype=torch.float64, device=self.device)
        index, nindex = (self._rys_groups(cls) if layout == self.RYS_GROUPED
                         else self._rys_strided(cls))
        # the LDS recursion pass reads the quartet-major layout, whose writes here are
        # contiguous across lanes
        lds = self._lds_hrr_plan(dev, cls, mode, rys=True)
        qmajor = lds is not None
        check(self.lib.gbx_eri_rys_acc(ctypes.byref(tab), ptr(self._rys_table()), int(layout),
                                       int(batch or 0), ptr(index), nindex, *cls,
                                       ptr(quartets), nq, ptr(count), R, ptr(acc),
                                       nacc if qmajor else nq, 1 if qmajor else 0,
                                       self._stream()),
              f"gbx_eri_rys_acc{cls}")
        self._hrr_pass(dev, tab, st, cls, quartets, count, mode, acc, R, out, qmajor=qmajor,
                       lds=lds)
        return "rys"

    LDS_BYTES = 160 * 1024   # LDS of one gfx950 workgroup
    HRR_SPLIT_LDS = 64 * 1024  # LDS bytes up to which gbx_eri_hrr_lds runs 256 or 512 lanes

    def _hrr_recipe(self, cls):
        """(host int32 recipe, device copy, LDS bytes of W per unit output dimension
        product, LDS bytes of one row buffer) of the LDS horizontal-recursion pass."""
        if cls not in self._recipes:
            from .programs import hrr_recipe, ncart

            la, lb, lc, ld = cls
            rec = np.ascontiguousarray(hrr_recipe(*cls), dtype=np.int32)
            na, nc, nlk, nlb = (int(x) for x in rec[:4])
            half = max([na, nc, ncart(lc) * ncart(ld), ncart(la) * ncart(lb)]
                       + [int(x) for x in rec[4:4 + nlk + nlb]])
            half += half & 1
            self._recipes[cls] = (rec, torch.from_numpy(rec).to(self.device), na * 8,
                                  2 * half * 8)
        return self._recipes[cls]

    def _lds_hrr_plan(self, dev, cls, mode, rys=False, split=False):
        """(host recipe, device recipe, largest C and D output dimensions) when the LDS
        horizontal-recursion pass runs this class and mode (``rys``: after the Rys
        primitive pass; ``split``: after the split register kernels), else None."""
        if self.hrr_lds == "never" or mode not in (GBX_MODE_PACKED, GBX_MODE_PACKED_ADD):
            return None
        rec, drec, w_unit, row = self._hrr_recipe(cls)
        dc = dev.max_out_dim.get(cls[2], 1)
        dd = dev.max_out_dim.get(cls[3], 1)
        if w_unit * dc * dd + row > self.LDS_BYTES:
            return None
        if (self.hrr_lds == "always" or rys
                or self._workspace(cls, GBX_FP64, 1, False)[1] > 0):
            return rec, drec, dc, dd
        w = w_unit * dc * dd  # the row groups and W, as gbx_eri_hrr_lds sizes them
        if split and min(16, (self.LDS_BYTES - w) // row) * row + w <= self.HRR_SPLIT_LDS:
            return rec, drec, dc, dd
        return None

    def _hrr_pass(self, dev, tab, st, cls, quartets, count, mode, acc, R, out, qmajor=False,
                  lds=None):
        """Horizontal recursion, norms, transforms and placement of a class from R partial
        contracted accumulators per quartet: the LDS pass (gbx_eri_hrr_lds) where
        ``_lds_hrr_plan`` selects it, else the cooperative pass (gbx_eri_hrr).  ``qmajor``:
        the accumulators are in the quartet-major layout (LDS pass only)."""
        nq = quartets.shape[0]
        if lds is None:
            lds = self._lds_hrr_plan(dev, cls, mode)
        if lds is not None:
            rec, drec, dc, dd = lds
            stride = st.header[1] if qmajor else nq
            check(self.lib.gbx_eri_hrr_lds(ctypes.byref(tab), *cls, ptr(rec), ptr(drec),
                                           int(rec.size), dc, dd, ptr(quartets), nq,
                                           ptr(count), mode, ptr(acc), stride, R,
                                           1 if qmajor else 0, ptr(out), self._stream()),
                  f"gbx_eri_hrr_lds{cls}")
            return
        if qmajor:
            raise ValueError("the cooperative horizontal-recursion pass reads the "
                             "accumulator-major layout only")
        sc = self.scratch()
        check(self.lib.gbx_eri_hrr(ctypes.byref(tab), ctypes.byref(st), *cls,
                                   ptr(quartets), nq, ptr(count), mode, ptr(acc), nq, R,
                                   ptr(out), ptr(sc), sc.numel(), self._stream()),
              f"gbx_eri_hrr{cls}")

    def schwarz(self, dev: DeviceTables):
        """Bounds per global shell pair, device tensor (npair, 2).

        Column 0: Q = sqrt(max |(ij|ij)|), un-normalised as the reference's screener.
        Column 1: sqrt(max |contracted [e0|f0] of (ij|ij)|), the bound on the inputs of
        the horizontal recursion that selects the FP32 recursion bulk.
        """
        q = torch.zeros((dev.batch.npair, 2), dtype=torch.float64, device=self.device)
        for cls, nprimq, dq in dev.schwarz_lists:
            # Q needs no output transform: the register kernels serve every class they cover
            self.launch_class(dev, cls, dq, nprimq, GBX_FP64, GBX_MODE_SCHWARZ, q,
                              path=self.path)
        return q

    def plan_quartets(self, dev: DeviceTables, q_pair=None, threshold=0.0, fp32_cutoff=0.0):
        """Screened, bucketed sub-quartet lists enumerated on the host (the device
        queue builder's result, for checks).

        Returns a list of Bucket, phase by phase.  Quartets whose bound Q_ij Q_kl is
        below ``fp32_cutoff`` run with an FP32 recursion bulk; all others in FP64.
        """
        b = dev.batch
        if q_pair is not None:
            q_pair = np.asarray(q_pair).reshape(b.npair, -1)
        bras, kets = [], []
        for mi, m in enumerate(b.mols):
            po = b.mol_pair_off[mi]
            qm = None if q_pair is None else q_pair[po:po + m.npair, 0]
            bra, ket = m.quartets(qm, threshold)
            bras.append(bra.astype(np.int64) + po)
            kets.append(ket.astype(np.int64) + po)
        bra = np.concatenate(bras)
        ket = np.concatenate(kets)
        plans = []
        if fp32_cutoff > 0:
            if q_pair is None:
                raise ValueError("mixed precision needs the Schwarz bounds")
            po_ = self._orig_bounds(b, q_pair[:, -1])
            low = po_[bra] * po_[ket] < fp32_cutoff
            groups = [(GBX_FP32_BULK, low), (GBX_FP64, ~low)]
        else:
            groups = [(GBX_FP64, np.ones(bra.size, dtype=bool))]
        nprim = b.pair_info[:, 3].astype(np.int64)
        for prec, mask in groups:
            if not mask.any():
                continue
            for phase, xb, xk in b.sub_quartets(bra[mask], ket[mask]):
                for cls, nprimq, dense, quart in b.class_buckets(xb, xk):
                    qb, qk = quart[:, 0].astype(np.int64), quart[:, 1].astype(np.int64)
                    eris = int(b.unique_counts(qb, qk).sum()) if phase in STORE_PHASES else 0
                    flops = float(quartet_flops(cls, nprim[qb] * nprim[qk], pure=dense,
                                                contracted=phase in STORE_PHASES).sum())
                    plans.append(Bucket(cls, nprimq, prec, dense,
                                        torch.from_numpy(quart).to(self.device), quart.shape[0],
                                        eris, flops, phase=phase))
        return plans

    def pair_bounds(self, dev: DeviceTables, q_pair):
        """(screening bound, precision bound) per global pair, each of its original-shell
        pair (max over the expanded pairs of a generalised contraction); device tensors."""
        b = dev.batch
        q = q_pair[:, 0].contiguous()
        p = q_pair[:, 1].contiguous()
        if not b.generalized:
            return q, p
        qh = q_pair.cpu().numpy()
        return (torch.from_numpy(self._orig_bounds(b, qh[:, 0])).to(self.device),
                torch.from_numpy(self._orig_bounds(b, qh[:, 1])).to(self.device))

    def plan_device(self, dev: DeviceTables, bound=None, threshold=0.0, fp32_cutoff=0.0,
                    stats=True, device_counts=False):
        """Build the screened, bucketed work queue on the device.

        ``bound`` is the (screening, precision) pair of per-pair device bounds from
        pair_bounds(); required when screening or mixed precision is requested
        (mixed precision is on when ``fp32_cutoff`` > 0 or the tables carry
        per-molecule cutoffs).  Returns a list of Bucket (quartets are views of one
        device buffer cached on ``dev``).

        stats=True: a counting pass first (exact capacities, packed-slot and
        primitive-quartet counts per bucket for FLOP / ERI bookkeeping).
        stats=False: one pass into capacities computed on the host from the pair
        keys; the counts are then read back (device_counts=False) or left on the
        device for the launches to read (device_counts=True: no host round trip).
        """
        b = dev.batch
        mixed = fp32_cutoff > 0 or dev.mol_fp32_cutoff is not None
        if bound is None:
            if threshold > 0 or mixed:
                raise ValueError("screening / mixed precision need the Schwarz bounds")
            ones = torch.ones(b.npair, dtype=torch.float64, device=self.device)
            bound = (ones, ones)
        qb, pbd = bound
        nb = NUM_BUCKETS
        st = self._stream()
        mask = self.fp32_ok
        mcut = dev.mol_fp32_cutoff
        if stats:
            counts = torch.zeros(nb, dtype=torch.int32, device=self.device)
            eris = torch.zeros(nb, dtype=torch.int64, device=self.device)
            primq = torch.zeros(nb, dtype=torch.int64, device=self.device)
            check(self.lib.gbx_quartets_count(len(b.mols), ptr(dev.mol_cand_off),
                                              ptr(dev.mol_pair_off), ptr(dev.pair_key),
                                              ptr(dev.pair_sub), ptr(dev.pair_npf), ptr(qb),
                                              ptr(pbd),
                                              float(threshold), float(fp32_cutoff), ptr(mcut),
                                              ptr(mask), b.n_cand, ptr(counts), ptr(eris),
                                              ptr(primq), st), "gbx_quartets_count")
            cap = counts.cpu().numpy().astype(np.int64)
            layout = ("stats", cap.tobytes())
        else:
            # room for every unscreened candidate of each bucket (both precision
            # variants when mixed); the FP32 variants follow the current mask, so a
            # mixed layout is only valid for the mask version it was sized with
            layout = ("cap", mixed, len(self._masks) if mixed else 0)
            cap = b.bucket_capacity().copy()
            if mixed:
                # the FP32 variant of a bucket only where the class may run FP32
                cap = cap.reshape(-1, 2, 16)
                allow = self._fp32_ok_host[np.arange(cap.shape[0]) % self._fp32_ok_host.size]
                cap[:, 1, :] = cap[:, 0, :] * allow[:, None]
                cap = cap.reshape(-1)
        if layout not in dev.queue:
            offs = np.concatenate([[0], np.cumsum(cap)]).astype(np.int64)   # nb + 1 entries
            total = max(int(cap.sum()), 1)
            if self.stream_workspaces:
                quart = self.workspace("queue", 2 * total, torch.int32).view(total, 2)
                d_offs = self.workspace("queue_offs", nb + 1, torch.int64)
                d_offs.copy_(torch.from_numpy(offs).pin_memory(), non_blocking=True)
                cursor = self.workspace("queue_cursor", nb, torch.int32)
            else:
                quart = torch.empty((total, 2), dtype=torch.int32, device=self.device)
                d_offs = _h2d(offs, self.device)
                cursor = torch.zeros(nb, dtype=torch.int32, device=self.device)
            dev.queue[layout] = {"cap": cap, "offs": offs, "quart": quart, "d_offs": d_offs,
                                 "cursor": cursor}
        qc = dev.queue[layout]
        quart, d_offs, cursor, offs = qc["quart"], qc["d_offs"], qc["cursor"], qc["offs"]
        cursor.zero_()
        check(self.lib.gbx_quartets_fill(len(b.mols), ptr(dev.mol_cand_off), ptr(dev.mol_pair_off),
                                         ptr(dev.pair_key), ptr(dev.pair_sub), ptr(qb), ptr(pbd),
                                         float(threshold),
                                         float(fp32_cutoff), ptr(mcut), ptr(mask), b.n_cand,
                                         ptr(d_offs), ptr(cursor), ptr(quart), st),
              "gbx_quartets_fill")
        if stats or device_counts:
            ch = cap
        else:
            ch = cursor.cpu().numpy().astype(np.int64)
            if np.any(ch > cap):
                raise RuntimeError("work-queue bucket capacity exceeded (host capacity model "
                                   "out of step with the device queue builder)")
        eh = eris.cpu().numpy() if stats else None
        ph = primq.cpu().numpy() if stats else None
        plans = []
        keys = np.flatnonzero(ch)
        heads = {}
        for key in keys.tolist():
            heads.setdefault(key >> 4, []).append(key)
        for head in sorted(heads):
            ks = heads[head]
            phase, cls, dense, fp32 = decode_head(head)
            prec = GBX_FP32_BULK if fp32 else GBX_FP64
            fused = ((not dense) and self.path != "coop" and dev.potential == "coulomb"
                     and self.has_reg(cls))
            fpr, fco = class_flops(*cls, pure=dense) if stats else (0, 0)
            if phase not in STORE_PHASES:
                fco = 0   # a shell quartet's contracted part is counted with its first
                          # sub-quartet
            if fused:
                base = head << 4
                n = int(ch[ks].sum())
                flops = float(ph[ks].sum()) * fpr + n * fco if stats else 0.0
                eris_n = int(eh[ks].sum()) if stats else 0
                seg = (d_offs[base:base + 17], cursor[base:base + 16],
                       int(cap[base:base + 16].sum()))
                plans.append(Bucket(cls, 1 << (ks[-1] & 15), prec, dense, quart, n, eris_n, flops,
                                    seg=seg, key=("head", head), phase=phase))
                continue
            for key in ks:
                n = int(ch[key])
                view = quart[offs[key]:offs[key] + n]
                flops = float(ph[key]) * fpr + n * fco if stats else 0.0
                plans.append(Bucket(cls, 1 << (key & 15), prec, dense, view, n,
                                    int(eh[key]) if stats else 0, flops,
                                    count=cursor[key:key + 1] if device_counts else None,
                                    key=("key", key), phase=phase))
        return plans

    def check_queue(self, dev: DeviceTables):
        """Raise if the last fill of any of ``dev``'s work queues reserved more entries for
        a bucket than its capacity (the fill drops those quartets, so the step's result
        would miss them).  Reads the device cursors: call it outside timed or captured
        regions."""
        for layout, qc in dev.queue.items():
            over = np.flatnonzero(qc["cursor"].cpu().numpy().astype(np.int64) > qc["cap"])
            if over.size:
                k = int(over[0])
                raise RuntimeError(
                    f"work-queue bucket capacity exceeded in {over.size} bucket(s), e.g. "
                    f"key {k} ({decode_head(k >> 4)}, bin {k & 15}) of queue {layout[0]}")

    def capture(self, dev: DeviceTables, bound=None, threshold=0.0, fp32_cutoff=0.0,
                out=None, tuner=None):
        """Record one whole step -- the work-queue fill and every ERI launch -- into a
        HIP graph (torch.cuda.CUDAGraph).  Counts stay on the device, so the graph
        is valid for every replay on the same tables; replaying it recomputes the
        packed array ``out`` in place.  Returns (graph, out, plans)."""
        if out is None:
            out = torch.zeros(dev.batch.n_slots, dtype=torch.float64, device=self.device)
        # eager warm-up: allocates the queue buffers and per-stream workspaces
        plans = self.plan_device(dev, bound, threshold, fp32_cutoff, stats=False,
                                 device_counts=True)
        if tuner is not None:
            tuner.apply(plans)
        self.run(dev, plans, out)
        torch.cuda.synchronize(self.device)
        graph = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(self.device)
        side.wait_stream(torch.cuda.current_stream(self.device))
        with torch.cuda.stream(side):
            plans = self.plan_device(dev, bound, threshold, fp32_cutoff, stats=False,
                                     device_counts=True)
            if tuner is not None:
                tuner.apply(plans)
            self.run(dev, plans, out)
        torch.cuda.synchronize(self.device)
        with torch.cuda.graph(graph, stream=side):
            plans = self.plan_device(dev, bound, threshold, fp32_cutoff, stats=False,
                                     device_counts=True)
            if tuner is not None:
                tuner.apply(plans)
            # with launch_streams > 1 the side streams fork from and join back into the
            # capturing stream through events, so the graph holds the parallel branches
            self.run(dev, plans, out)
        return graph, out, plans

    @staticmethod
    def _orig_bounds(b, q_pair):
        """Per global pair: bound of its original-shell pair (max over expanded pairs)."""
        out = np.empty(b.npair)
        for mi, m in enumerate(b.mols):
            po = b.mol_pair_off[mi]
            Qo = m.schwarz_orig(q_pair[po:po + m.npair])
            oi, oj = m.pair_orig
            out[po:po + m.npair] = Qo[oi, oj]
        return out

    def _side_streams(self, n):
        key = torch.cuda.current_stream(self.device).cuda_stream
        lst = self._streams.setdefault(key, [])
        while len(lst) < n:
            lst.append(torch.cuda.Stream(self.device))   # default priority
        return lst[:n]

    def run(self, dev: DeviceTables, plans, out=None, n_streams=None):
        """Execute the plans into the packed array (allocated if None).

        Phase 0 stores every packed slot of every kept quartet; phases 1-3 add to
        them, one phase after the other.  Within a phase every packed slot belongs to
        one sub-quartet, so with ``n_streams`` > 1 (default: ``self.launch_streams``)
        a phase's buckets are spread round-robin over that many side streams forked
        from and joined back into the current stream: small launches of different
        classes overlap on the device instead of running one after another.  (Giving
        the quartets with a split pair a phase of their own, so that their later
        phases could overlap the rest, doubles the launches of the heaviest classes
        and measured slower.)"""
        if out is None:
            out = torch.zeros(dev.batch.n_slots, dtype=torch.float64, device=self.device)
        ns = self.launch_streams if n_streams is None else n_streams
        phases = sorted({bk.phase for bk in plans})
        for ph in phases:
            group = [bk for bk in plans if bk.phase == ph]
            if ns <= 1 or len(group) <= 1:
                for bk in group:
                    bk.path = self.launch_bucket(dev, bk, out)
                continue
            cur = torch.cuda.current_stream(self.device)
            side = self._side_streams(ns)
            fork = torch.cuda.Event()
            fork.record(cur)
            for st in side:
                st.wait_event(fork)
            for i, bk in enumerate(group):
                with torch.cuda.stream(side[i % ns]):
                    bk.path = self.launch_bucket(dev, bk, out)
            for st in side:
                join = torch.cuda.Event()
                join.record(st)
                cur.wait_event(join)
        return out

    FP32_MAX_L = 2

    def set_fp32_allowed(self, allowed):
        """allowed: {(cls, dense): bool}; classes not listed keep the default (FP32 allowed
        iff every l <= FP32_MAX_L).  A listed class is never allowed beyond the default."""
        c = np.arange(8 ** 4)
        lmax = np.maximum.reduce([(c >> 9) & 7, (c >> 6) & 7, (c >> 3) & 7, c & 7])
        mask = np.repeat((lmax <= self.FP32_MAX_L).astype(np.int8), 2)
        for (cls, dense), ok in allowed.items():
            ci = ((cls[0] * 8 + cls[1]) * 8 + cls[2]) * 8 + cls[3]
            if not ok:
                mask[ci * 2 + int(dense)] = 0
        host = mask.astype(np.int64)
        if self._masks and np.array_equal(host, self._fp32_ok_host):
            return
        self._fp32_ok_host = host
        self.fp32_ok = torch.from_numpy(mask).to(self.device)
        self._masks.append(self.fp32_ok)

    def unpack_dense(self, packed, nfunc, notation):
        dense = torch.empty((nfunc,) * 4, dtype=torch.float64, device=self.device)
        check(self.lib.gbx_unpack_dense(ptr(packed), nfunc, 1 if notation == "physicist" else 0,
                                        ptr(dense), self._stream()), "gbx_unpack_dense")
        return dense

    def mo_transform(self, packed, nfunc, transform, notation):
        """Dense (m, m, m, m) device array of sum_ijkl T_ai T_bj T_ck T_dl (ij|kl) from the
        packed ERIs (``transform``: m x nfunc), in chemist or physicist order; the
        workspace (the half-transformed nfunc(nfunc+1)/2 x m(m+1)/2 array and one
        chunk of GEMM staging) is the only intermediate."""
        tr = torch.as_tensor(np.ascontiguousarray(transform, dtype=np.float64),
                             device=self.device)
        if tr.ndim != 2 or tr.shape[1] != nfunc:
            raise ValueError(f"transform must have shape (m, {nfunc}), got {tuple(tr.shape)}")
        m = int(tr.shape[0])
        trt = tr.t().contiguous()
        n_work = ctypes.c_int64()
        check(self.lib.gbx_mo_workspace(nfunc, m, ctypes.byref(n_work)), "gbx_mo_workspace")
        work = torch.empty(n_work.value, dtype=torch.float64, device=self.device)
        out = torch.empty((m,) * 4, dtype=torch.float64, device=self.device)
        check(self.lib.gbx_mo_transform(ptr(packed), nfunc, ptr(tr), ptr(trt), m, ptr(work),
                                        ptr(out), 1 if notation == "physicist" else 0,
                                        self._stream()), "gbx_mo_transform")
        return out

    def jk(self, packed, nfunc, density):
        """Coulomb and exchange matrices (device float64, n x n) of one molecule's
        packed ERIs for a symmetric density."""
        n = int(nfunc)
        npair = n * (n + 1) // 2
        if packed.numel() != npair * (npair + 1) // 2:
            raise ValueError(f"packed array has {packed.numel()} slots, not the "
                             f"{npair * (npair + 1) // 2} of {n} functions")
        D = torch.as_tensor(density, dtype=torch.float64, device=self.device).contiguous()
        if D.shape != (n, n):
            raise ValueError(f"density must be ({n}, {n}), got {tuple(D.shape)}")
        J = torch.empty((n, n), dtype=torch.float64, device=self.device)
        K = torch.empty((n, n), dtype=torch.float64, device=self.device)
        check(self.lib.gbx_jk_packed(ptr(packed), n, ptr(D), ptr(J), ptr(K), self._stream()),
              "gbx_jk_packed")
        return J, K

    def boys(self, dev, T, mmax):
        """F_m(T), m = 0..mmax, as a device (n, mmax + 1) tensor.  ``dev``: the tables
        whose Boys table to use, or None for the engine's own copy."""
        T = torch.as_tensor(T, dtype=torch.float64, device=self.device).contiguous()
        out = torch.empty((T.numel(), mmax + 1), dtype=torch.float64, device=self.device)
        st = dev.struct if dev is not None else self._boys_struct()
        check(self.lib.gbx_boys(ctypes.byref(st), ptr(T), T.numel(), mmax, ptr(out),
                                self._stream()), "gbx_boys")
        return out

    def _boys_struct(self):
        """Tables struct with only the Boys table set (for standalone Boys values)."""
        if getattr(self, "_boys_tab", None) is None:
            tab = boys_table()
            self._boys_dev = torch.tensor(np.asarray(tab), dtype=torch.float64,
                                         device=self.device)
            st = GbxTables()
            st.boys_table = ptr(self._boys_dev)
            st.boys_nt = tab.shape[0]
            st.boys_nord = tab.shape[1]
            st.boys_h = BOYS_H
            self._boys_tab = st
        return self._boys_tab

    @property
    def boys_max_order(self):
        """Largest order the device Boys evaluation serves (table columns less the
        Taylor terms)."""
        return boys_table().shape[1] - 1 - 8


_ENGINES = {}


def get_engine(device="cuda"):
    key = str(torch.device(device))
    if key not in _ENGINES:
        _ENGINES[key] = EriEngine(device)
    return _ENGINES[key]


def molecule_tables(basis, coord_types=None):
    return MoleculeTables(basis, coord_types)
