"""The autotuner and the wave-to-bucket schedules never change results: the
hit set after autotuning equals the oracle's."""
import numpy as np
import pytest

import _oracle as orc
from gpusharesat_amd import workloads
from gpusharesat_amd.engine import ClauseMap, Engine, TUNE_SCHEDULE

pytestmark = pytest.mark.gpu


def records(e, lens, rows):
    hits = e.hits()
    ids = ClauseMap(lens, rows).ids(hits)
    out = np.empty(hits.size, dtype=orc.HIT_DTYPE)
    out["clause"] = ids
    out["word"] = hits["word"]
    out["false_mask"] = hits["false_mask"]
    out["unit_mask"] = hits["unit_mask"]
    return out[np.lexsort((out["word"], out["clause"]))]


def test_autotune_and_schedules_keep_parity():
    inst = workloads.sweep_1024(seed=9, scale=0.05)
    exp = orc.sweep(inst.offsets, inst.lits, inst.is_def, inst.is_true, inst.valid)
    with Engine(0) as e:
        rows = e.append(inst.lits, inst.lens)
        e.reset_assignments(inst.var_count, inst.words)
        e.upload(inst.is_def, inst.is_true)
        for sched in (0, 1, 2):
            e.set_tuning(TUNE_SCHEDULE, sched)
            e.sweep(inst.valid)
            got = records(e, inst.lens, rows)
            assert got.tobytes() == exp.tobytes()
        best, ms = e.autotune(inst.valid, reps=1)
        print("autotuned:", best, ms, flush=True)
        assert ms > 0 and best["unroll"] in (2, 4) and best["schedule"] in (1, 2)
        e.sweep(inst.valid)
        got = records(e, inst.lens, rows)
        assert got.tobytes() == exp.tobytes()
