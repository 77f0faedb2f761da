"""gbx_eri_coop_noalloc (the entry point that allocates nothing) against gbx_eri_coop,
both called through the C ABI: bit-identical packed output on two calls in a row, the
second with fewer quartets, for a class whose workspace exceeds the LDS (it runs in the
library's file-scope scratch array) and one that fits the LDS."""
import ctypes

import pytest

from tests.molecules import random_basis

pytestmark = pytest.mark.gpu


def _class_quartets(eng, dev, cls):
    """Phase-0 (bra, ket) entries of every primitive bin of one class, device int32."""
    import torch

    parts = [bk.quart for bk in eng.plan_quartets(dev)
             if bk.phase == 0 and tuple(bk.cls) == cls and not bk.dense]
    assert parts, f"no phase-0 bucket of class {cls}"
    return torch.cat(parts).contiguous()


@pytest.mark.parametrize("ls,cls", [([4, 4, 4], (4, 4, 4, 4)), ([2, 1, 2, 1], (2, 1, 2, 1))])
def test_noalloc_matches_coop(gpu, ls, cls):
    import torch

    from gbasis_amd._lib import GBX_FP64, GBX_MODE_PACKED, check, ptr
    from gbasis_amd.engine import DeviceTables, get_engine
    from gbasis_amd.tables import BatchTables, MoleculeTables

    print(f"[gpu case] no-alloc entry point, class {cls}", flush=True)
    eng = get_engine()
    basis = random_basis(61 + sum(ls), ls, kmin=1, kmax=2, emin=0.3, emax=5.0)
    dev = DeviceTables(BatchTables([MoleculeTables(basis)]))
    st, _, _ = eng.programs.get(cls)
    lds, glb = ctypes.c_int64(), ctypes.c_int64()
    check(eng.lib.gbx_coop_workspace(ctypes.byref(st), *cls, GBX_FP64, 1, 0, ctypes.byref(lds),
                                     ctypes.byref(glb)), "gbx_coop_workspace")
    uses_scratch = glb.value > 0
    assert uses_scratch == (cls == (4, 4, 4, 4)), (cls, lds.value, glb.value)
    sc = eng.scratch()
    quart = _class_quartets(eng, dev, cls)
    n_all = quart.shape[0]
    assert n_all >= 4
    for n in (n_all, max(1, n_all // 3)):
        q = quart[:n].contiguous()
        ref = torch.zeros(dev.batch.n_slots, dtype=torch.float64, device="cuda")
        got = torch.zeros_like(ref)
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        check(eng.lib.gbx_eri_coop(ctypes.byref(dev.struct), ctypes.byref(st), *cls, ptr(q), n,
                                   None, GBX_FP64, 1, GBX_MODE_PACKED, ptr(ref), ptr(sc),
                                   sc.numel(), stream), "gbx_eri_coop")
        before = torch.cuda.memory_allocated()
        check(eng.lib.gbx_eri_coop_noalloc(ctypes.byref(dev.struct), ctypes.byref(st), *cls,
                                           ptr(q), n, None, GBX_FP64, 1, GBX_MODE_PACKED,
                                           ptr(got), stream), "gbx_eri_coop_noalloc")
        torch.cuda.synchronize()
        assert torch.cuda.memory_allocated() == before
        assert bool(ref.abs().max() > 0)
        assert torch.equal(got, ref), float((got - ref).abs().max())
