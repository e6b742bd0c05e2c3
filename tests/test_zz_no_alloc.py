"""gss_hits_gather, the entry point that allocates nothing, through the C
ABI: the records it writes into caller device memory are exactly those
gss_hits_copy returns for the same sweep (and the oracle's), on two calls in
a row on one engine, the second on a smaller DB and table.  The device
buffer comes from the HIP runtime the library itself uses (hipMalloc through
ctypes), so no other allocator is involved."""
import ctypes

import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd import _lib, workloads
from gpusharesat_amd.engine import HIT_DTYPE, ClauseMap, Engine

pytestmark = pytest.mark.gpu


def _hip():
    _lib.load()  # the library's own runtime is then already loaded
    hip = ctypes.CDLL("libamdhip64.so.7")
    hip.hipMalloc.argtypes = [ctypes.c_void_p, ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return hip


def _gather(e: Engine, hip, cap: int) -> np.ndarray:
    dev = ctypes.c_void_p()
    assert hip.hipMalloc(ctypes.byref(dev), max(1, cap) * HIT_DTYPE.itemsize) == 0
    try:
        n = ctypes.c_int64()
        _lib.call("gss_hits_gather", e.handle, dev, cap, ctypes.byref(n))
        out = np.empty(n.value, HIT_DTYPE)
        if n.value:
            assert hip.hipMemcpy(out.ctypes.data, dev, n.value * HIT_DTYPE.itemsize, 2) == 0
        return out
    finally:
        hip.hipFree(dev)


def _key(h: np.ndarray) -> np.ndarray:
    return np.sort(h, order=["len", "row", "word"])


def _check(e: Engine, hip, inst, rows):
    print(f"case: {inst.lens.size} clauses, {inst.var_count} vars", flush=True)
    copied = e.hits()
    n = ctypes.c_int64()
    _lib.call("gss_hits_gather", e.handle, None, 0, ctypes.byref(n))  # count only
    assert n.value == copied.size
    gathered = _gather(e, hip, copied.size)
    assert np.array_equal(_key(gathered), _key(copied))
    ids = ClauseMap(inst.lens, rows).ids(gathered)
    got = sorted(zip(ids.tolist(), gathered["word"].tolist(), gathered["false_mask"].tolist(),
                     gathered["unit_mask"].tolist()))
    exp = orc.sweep(inst.offsets, inst.lits, inst.is_def, inst.is_true, inst.valid)
    assert got == sorted(zip(exp["clause"].tolist(), exp["word"].tolist(),
                             exp["false_mask"].tolist(), exp["unit_mask"].tolist()))
    return copied.size


def test_gather_matches_copy_twice_second_smaller():
    hip = _hip()
    big = workloads.plain_64(seed=11, scale=0.3)
    small = workloads.plain_64(seed=12, scale=0.05)
    with Engine(0) as e:
        rows = e.append(big.lits, big.lens)
        e.reset_assignments(big.var_count, big.words)
        e.upload(big.is_def, big.is_true)
        e.sweep(big.valid)
        n1 = _check(e, hip, big, rows)
        # too small an output is refused before anything is written
        from gpusharesat_amd._lib import GssError
        with pytest.raises(GssError, match="output holds"):
            _gather(e, hip, n1 - 1)
        e.clear()
        rows = e.append(small.lits, small.lens)
        e.reset_assignments(small.var_count, small.words)
        e.upload(small.is_def, small.is_true)
        e.sweep(small.valid)
        n2 = _check(e, hip, small, rows)
        assert 0 < n2 < n1
