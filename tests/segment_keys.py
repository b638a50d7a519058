"""Keys whose Fibonacci hashes (key * 0x9E3779B9 mod 2^32, the engine's table hash) all fall into
the first segment of a pipeline table: for 2048 < n <= 4096 build tuples the table has 2^13 key
slots in 2048 buckets of 4, linear probing wraps within segments of 512 buckets (2048 slots), and
a hash's segment is its top 2 bits."""
import numpy as np

FIB = 0x9E3779B9
FIB_INV = pow(FIB, -1, 1 << 32)
SEG_SLOTS = 2048


def key_of_hash(h):
    """the INT32 key whose hash is h"""
    k = (np.asarray(h, dtype=np.uint64) * FIB_INV) % (1 << 32)
    return k.astype(np.uint32).view(np.int32)


def hash_of(keys):
    return (np.asarray(keys, dtype=np.int64).astype(np.uint64) % (1 << 32) * FIB) % (1 << 32)


def segment_keys(rng, in_seg0, elsewhere):
    """in_seg0 distinct keys hashing into segment 0 (top 2 hash bits 00) and `elsewhere` keys
    hashing into segments 1..3; returns (keys, spare) where spare are more distinct segment-0
    keys that are not in keys (probe keys that miss)"""
    h0 = rng.choice(1 << 30, size=in_seg0 + 500, replace=False).astype(np.uint64)
    h1 = ((1 << 30) + rng.choice(3 << 30, size=elsewhere, replace=False)).astype(np.uint64)
    keys = np.concatenate([key_of_hash(h0[:in_seg0]), key_of_hash(h1)])
    return keys, key_of_hash(h0[in_seg0:])
