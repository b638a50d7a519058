"""A model, on the CPU, of a pipeline table segment being filled (the segment pass's LDS insert
and the compare-and-swap insert both probe the buckets of one 512-bucket segment, wrapping, and
flag the build when a key finds no free slot) and probed (a lookup that finds neither its key nor
a free slot stops after the segment's 512 buckets), over the crafted keys that the GPU test
tests/test_gpu_segment_full.py uses."""
import numpy as np

from segment_keys import SEG_SLOTS, hash_of, segment_keys

K = 13  # key slots 2^13 for 2048 < n <= 4096 tuples


def model_fill(keys):
    """(table, full flag): segment-local linear probing over 4-key buckets, as the engine's
    inserts do; the table maps bucket -> list of keys"""
    buckets = {}
    full = False
    for key in keys:
        h = int(hash_of([key])[0])
        s0 = h >> (32 - K)
        b = s0 >> 2
        seg, smask = b >> 9, 511
        placed = False
        for m in range(smask + 1):  # bounded: the segment's buckets
            bb = (seg << 9) | ((b + m) & smask)
            slot = buckets.setdefault(bb, [])
            if key in slot:
                full = True  # a repeated key
                placed = True
                break
            if len(slot) < 4:
                slot.append(key)
                placed = True
                break
        if not placed:
            full = True
    return buckets, full


def model_lookup(buckets, key):
    """(found, buckets read)"""
    h = int(hash_of([key])[0])
    b = (h >> (32 - K)) >> 2
    seg, smask = b >> 9, 511
    for m in range(smask + 1):
        slot = buckets.get((seg << 9) | ((b + m) & smask), [])
        if key in slot:
            return True, m + 1
        if len(slot) < 4:
            return False, m + 1
    return False, smask + 1


def test_keys_fall_into_segment_zero():
    rng = np.random.default_rng(1)
    keys, spare = segment_keys(rng, SEG_SLOTS, 100)
    h = hash_of(keys)
    assert ((h[:SEG_SLOTS] >> 30) == 0).all() and ((h[SEG_SLOTS:] >> 30) != 0).all()
    assert ((hash_of(spare) >> 30) == 0).all()
    assert len(set(keys.tolist())) == len(keys) and not set(spare.tolist()) & set(keys.tolist())
    assert 2048 < len(keys) <= 4096  # a table of 2^13 key slots


def test_exactly_full_segment_is_no_overflow_and_lookups_end():
    rng = np.random.default_rng(2)
    keys, spare = segment_keys(rng, SEG_SLOTS, 100)
    buckets, full = model_fill(keys)
    assert not full
    seg0 = sum(len(v) for b, v in buckets.items() if b >> 9 == 0)
    assert seg0 == SEG_SLOTS  # every slot of segment 0 taken
    for k in keys[:200]:
        assert model_lookup(buckets, k)[0]
    for k in spare[:50]:  # an absent key in a full segment: all 512 buckets, then not found
        assert model_lookup(buckets, k) == (False, 512)


def test_one_key_too_many_flags_the_build():
    rng = np.random.default_rng(3)
    keys, _ = segment_keys(rng, SEG_SLOTS + 1, 100)
    _, full = model_fill(keys)
    assert full
