"""The autotuner's selection and cache logic (python/autotune.py) on the CPU:
the device calls are replaced by a stand-in engine whose speed depends on
the candidate, so the tuner must find the known best, leave the engine set to
it, and reuse the cached pick without measuring again."""
import os
import sys

import pytest

from conftest import PKG

sys.path.insert(0, os.path.join(PKG, "python"))


class FakeEngine:
    """rans_sst_py with the engine calls replaced: evaluate_many reports every
    member's cap and takes a time set by the current (block, slice)."""

    def __init__(self, rs, best=(256, 200)):
        self._rs = rs
        self.block, self.slice_us, self.calls = 0, 0, 0
        self.best = best

    def __getattr__(self, name):
        return getattr(self._rs, name)

    def engine_device_info(self):
        return {"cu_count": 256, "wavefront_size": 64}

    def engine_set_launch(self, iters, block):
        self.block = block

    def engine_set_time_slice(self, us):
        self.slice_us = us

    def evaluate_many(self, fms, thetas):
        import time
        self.calls += 1
        slow = 0.002 if (self.block, self.slice_us // 1000) == self.best else 0.02
        time.sleep(slow)

        class R:
            simple_iters = 50
        return [R() for _ in fms]


@pytest.fixture(scope="module")
def rs_mod():
    import rans_sst_py
    return rans_sst_py


def test_member_tuner_picks_fastest_and_caches(rs_mod, tmp_path):
    import autotune
    import ensembles
    eng = FakeEngine(rs_mod)
    specs = ensembles.channel_ensemble(4, 0)
    cache = str(tmp_path / "tune.json")
    choice, table = autotune.tune_members(eng, specs, cache_path=cache, sample_iters=50)
    assert (choice["block"], choice["slice_ms"]) == (256, 200)
    assert len(table) == len(autotune.MEMBER_CANDIDATES) and eng.calls == len(table)
    assert (eng.block, eng.slice_us) == (256, 200000)
    # second call: from the cache, no measurement, engine set to the pick
    eng2 = FakeEngine(rs_mod, best=(512, 100))
    choice2, _ = autotune.tune_members(eng2, specs, cache_path=cache, sample_iters=50)
    assert eng2.calls == 0 and (choice2["block"], choice2["slice_ms"]) == (256, 200)
    assert (eng2.block, eng2.slice_us) == (256, 200000)
    # another workload signature measures again
    choice3, _ = autotune.tune_members(eng2, ensembles.channel_ensemble(5, 1), cache_path=cache, sample_iters=50)
    assert eng2.calls == len(autotune.MEMBER_CANDIDATES) and (choice3["block"], choice3["slice_ms"]) == (512, 100)
