"""Hash structures of the oracle against the reference headers compiled in place (oracle/_ref):
PartitionedDirectoryHash (the table execute() uses, include/partitioned_directory_hash.h) and
UnchainedHash (include/unchained.h), plus the reference's tests/unchained_tests.cpp cases."""
import ctypes as C

import numpy as np
import pytest


def pairs_sorted(buf, n):
    p = buf[:2 * n].reshape(-1, 2)
    return p[np.lexsort((p[:, 1], p[:, 0]))]


def join_pairs(fn, bk, bv, pk, pv, threads):
    cap = 4 * (len(bk) * 4 + len(pk) * 4 + 16)
    out = np.zeros(2 * cap, dtype=np.uint64)
    n = fn(bk.ctypes.data, bv.ctypes.data, len(bk), pk.ctypes.data, pv.ctypes.data, len(pk),
           threads, out.ctypes.data, cap)
    assert n != 2**64 - 1
    return pairs_sorted(out, n)


@pytest.mark.parametrize("seed,threads", [(0, 2), (1, 3), (2, 4), (3, 8), (4, 5)])
def test_pdh_join_matches_reference(oracle, ref, seed, threads):
    rng = np.random.default_rng(seed)
    nb, np_ = int(rng.integers(0, 5000)), int(rng.integers(0, 8000))
    dom = int(rng.integers(1, 3000))
    bk = rng.integers(0, dom, nb).astype(np.uint32)
    pk = rng.integers(0, dom, np_).astype(np.uint32)
    bv = (rng.random(nb) > 0.1).astype(np.uint8)
    pv = (rng.random(np_) > 0.1).astype(np.uint8)
    a = join_pairs(oracle.lib.oracle_pdh_join, bk, bv, pk, pv, threads)
    b = join_pairs(ref.ref_pdh_join, bk, bv, pk, pv, threads)
    assert np.array_equal(a, b)
    # and the textbook equi-join: (probe i, build j) for every equal non-NULL key pair
    want = [(i, j) for i in range(np_) if pv[i] for j in np.nonzero((bk == pk[i]) & (bv == 1))[0]]
    assert len(a) == len(want)


def test_pdh_join_full_u32_keys(oracle, ref):
    rng = np.random.default_rng(42)
    bk = rng.integers(0, 2**32, 20000, dtype=np.uint64).astype(np.uint32)
    pk = np.concatenate([bk[::3], rng.integers(0, 2**32, 5000, dtype=np.uint64).astype(np.uint32)])
    one_b, one_p = np.ones(len(bk), np.uint8), np.ones(len(pk), np.uint8)
    assert np.array_equal(join_pairs(oracle.lib.oracle_pdh_join, bk, one_b, pk, one_p, 4),
                          join_pairs(ref.ref_pdh_join, bk, one_b, pk, one_p, 4))


class Unchained:
    def __init__(self, oracle):
        self.o = oracle.lib
        self.h = self.o.oracle_unchained_new()

    def insert(self, k, p):
        self.o.oracle_unchained_insert(self.h, k, p)

    def finalize(self):
        self.o.oracle_unchained_finalize(self.h)

    def probe(self, k):
        buf = (C.c_uint64 * 64)()
        n = self.o.oracle_unchained_probe(self.h, k, buf, 64)
        return sorted(buf[i] for i in range(min(n, 64)))

    def clear(self):
        self.o.oracle_unchained_clear(self.h)

    def __del__(self):
        self.o.oracle_unchained_free(self.h)


def test_unchained_reference_cases(oracle):
    # tests/unchained_tests.cpp:10-184 (integer-key cases)
    h = Unchained(oracle)
    assert h.probe(42) == []                                  # :10 probe on empty
    h.finalize()
    assert h.probe(1) == []                                   # :19 finalize on empty
    h.insert(1, 111)
    assert h.probe(1) == []                                   # :28 probe before finalize
    h.clear()
    h.insert(7, 123)
    h.finalize()
    assert h.probe(7) == [123] and h.probe(8) == []           # :38, :52
    h = Unchained(oracle)
    for i in range(10000):
        h.insert(i, i * 10)
    h.finalize()
    for k in (0, 1, 2, 17, 999, 4096, 9999):                 # :74
        assert h.probe(k) == [k * 10]
    h = Unchained(oracle)
    for p in (10, 20, 30):
        h.insert(5, p)
    h.finalize()
    assert h.probe(5) == [10, 20, 30]                         # :93 duplicates
    for n in (1, 2, 3, 4, 5, 7, 8, 9, 15, 16, 17):            # :159 power-of-two boundaries
        h = Unchained(oracle)
        for i in range(n):
            h.insert(i, i + 100)
        h.finalize()
        assert oracle.lib.oracle_unchained_dir_size(h.h) == 1 << (n - 1).bit_length()
        for k in (0, n // 2, n - 1):
            assert h.probe(k) == [k + 100]


@pytest.mark.parametrize("seed", range(4))
def test_unchained_matches_reference(oracle, ref, seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(0, 3000))
    keys = rng.integers(0, max(1, n // 2), n).astype(np.uint64)
    pay = rng.integers(0, 2**40, n).astype(np.uint64)
    probes = rng.integers(0, max(1, n), 2000).astype(np.uint64)
    counts = np.zeros(len(probes), np.uint64)
    out = np.zeros(len(probes) * 64 + 64, np.uint64)
    ref.ref_unchained(keys.ctypes.data, pay.ctypes.data, n, 1, probes.ctypes.data, len(probes),
                      counts.ctypes.data, out.ctypes.data, len(out))
    h = Unchained(oracle)
    for k, p in zip(keys.tolist(), pay.tolist()):
        h.insert(k, p)
    h.finalize()
    pos = 0
    for j, k in enumerate(probes.tolist()):
        c = int(counts[j])
        assert h.probe(k) == sorted(out[pos:pos + c].tolist())
        pos += c
