"""A pipeline table whose first segment is exactly full (every lookup of an absent key hashing
there reads all 512 buckets of the segment, then stops) and one whose first segment overflows
(the build is flagged and the query runs join-at-a-time instead), in both build modes: results
equal to the oracle's. The keys are crafted from the table hash (tests/segment_keys.py); the CPU
model of the same fill is tests/test_segment_full.py."""
import numpy as np
import pytest

import hjtest
from segment_keys import SEG_SLOTS, segment_keys

pytestmark = pytest.mark.gpu
I32 = hjtest.DataType.INT32


def _plan(oracle, in_seg0, seed):
    rng = np.random.default_rng(seed)
    bkeys, spare = segment_keys(rng, in_seg0, 100)
    nb = len(bkeys)
    probe = np.concatenate([rng.choice(bkeys, 30_000), rng.choice(spare, 10_000)])
    rng.shuffle(probe)
    pv = (rng.random(len(probe)) > 0.05).astype(np.uint8)
    plan = hjtest.Plan()
    plan.inputs.append(hjtest.ColumnarTable(len(probe), [
        oracle.codec.encode_int32(probe, pv),
        oracle.codec.encode_int32(np.arange(len(probe), dtype=np.int32))]))
    plan.inputs.append(hjtest.ColumnarTable(nb, [
        oracle.codec.encode_int32(bkeys),
        oracle.codec.encode_int32(np.arange(nb, dtype=np.int32))]))
    s0 = plan.new_scan_node(0, [(0, I32), (1, I32)])
    s1 = plan.new_scan_node(1, [(0, I32), (1, I32)])
    plan.root = plan.new_join_node(False, s0, s1, 0, 0, [(0, I32), (1, I32), (3, I32)])
    return plan


@pytest.mark.parametrize("in_seg0", [SEG_SLOTS, SEG_SLOTS + 1], ids=["full", "overflow"])
@pytest.mark.parametrize("part", [1, 0], ids=["segment-pass", "cas-insert"])
def test_full_segment(hjo, gpu_ctx, oracle, in_seg0, part):
    plan = _plan(oracle, in_seg0, 90 + in_seg0 % 7)
    want = oracle.rows(oracle.execute(plan))
    assert len(want) > 20_000
    print(f"segment keys {in_seg0}, partitioned build {part}", flush=True)
    gpu_ctx.set_option("partitioned_build", part)
    try:
        got = oracle.rows(gpu_ctx.execute(plan))
    finally:
        gpu_ctx.set_option("partitioned_build", 1)
    assert got == want
