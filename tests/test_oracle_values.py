"""value_t / StrRef bit layout: the reference's tests/valuet_tests.cpp cases, and the oracle
against the reference header include/value_t.h compiled in place (oracle/_ref)."""
import ctypes as C

import numpy as np
import pytest

import hjtest


def strref(oracle, *a):
    out = C.c_uint64()
    rc = oracle.lib.oracle_strref_encode(*a, C.byref(out))
    return rc, out.value


def fields(r):
    return ((r >> 51) & 1023, (r >> 41) & 1023, (r >> 23) & ((1 << 18) - 1),
            (r >> 7) & 0xFFFF, bool(r >> 61 & 1))


def test_valuet_strref_basic(oracle):
    # valuet_tests.cpp:8-16
    rc, r = strref(oracle, 0, 5, 7, 9, 11)
    assert rc == 0 and fields(r) == (5, 7, 9, 11, False)


def test_valuet_strref_long(oracle):
    # valuet_tests.cpp:18-26
    rc, r = strref(oracle, 1, 1, 2, 3, 4)
    assert rc == 0 and fields(r) == (1, 2, 3, 4, True)


def test_valuet_int32(oracle):
    # valuet_tests.cpp:50-69: tag 01 and the low 32 bits, negative values too
    for v in (12345, -999, 0, 2**31 - 1, -2**31):
        x = oracle.lib.oracle_value_int32(v)
        assert x >> 62 == 1
        assert C.c_int32(x & 0xFFFFFFFF).value == v


def test_valuet_strref_tag(oracle):
    # valuet_tests.cpp:71-90
    rc, r = strref(oracle, 1, 2, 4, 6, 8)
    v = oracle.lib.oracle_value_strref(r)
    assert v >> 62 == 2
    assert v & ~(3 << 62) == r & ~(3 << 62)


def test_strref_overflow(oracle):
    # value_t.h:36-47 throws std::out_of_range on any field overflow
    assert strref(oracle, 0, 1024, 0, 0, 0)[0] != 0
    assert strref(oracle, 0, 0, 1024, 0, 0)[0] != 0
    assert strref(oracle, 0, 0, 0, 1 << 18, 0)[0] != 0
    assert strref(oracle, 0, 0, 0, 0, 1 << 16)[0] != 0
    assert strref(oracle, 0, 1023, 1023, (1 << 18) - 1, 65535)[0] == 0


def test_splitmix64_known_values(oracle):
    # SplitMix64 (Steele, Lea, Flood 2014): the first outputs of the generator seeded with 0
    # are mix(0 + k*golden) for k = 1, 2, ...; splitmix64(x) = mix(x + golden)
    g = 0x9E3779B97F4A7C15
    assert oracle.lib.oracle_splitmix64(0) == 0xE220A8397B1DCDAF
    assert oracle.lib.oracle_splitmix64(g) == 0x6E789E6AA1B965F4
    assert oracle.lib.oracle_splitmix64((2 * g) % 2**64) == 0x06C45D188009454F


def test_against_reference_value_t(oracle, ref):
    rng = np.random.default_rng(5)
    for _ in range(2000):
        t, c, p, o = (int(rng.integers(0, 1024)), int(rng.integers(0, 1024)),
                      int(rng.integers(0, 1 << 18)), int(rng.integers(0, 1 << 16)))
        lg = int(rng.integers(0, 2))
        err = C.c_int()
        want = ref.ref_strref(lg, t, c, p, o, C.byref(err))
        rc, got = strref(oracle, lg, t, c, p, o)
        assert err.value == 0 and rc == 0 and got == want
        assert oracle.lib.oracle_value_strref(got) == ref.ref_value_strref(want)
        v = int(rng.integers(-2**31, 2**31))
        assert oracle.lib.oracle_value_int32(v) == ref.ref_value_int32(v)
    for bad in ((0, 1024, 0, 0, 0), (0, 0, 0, 0, 1 << 16)):
        err = C.c_int()
        ref.ref_strref(*bad, C.byref(err))
        assert err.value != 0 and strref(oracle, *bad)[0] != 0


def test_against_reference_splitmix(oracle, ref):
    rng = np.random.default_rng(9)
    for x in rng.integers(0, 2**63, size=1000, dtype=np.uint64).tolist() + [0, 1, 2**64 - 1]:
        assert oracle.lib.oracle_splitmix64(x) == ref.ref_splitmix64(x)


def ref_message(ref, args):
    buf = C.create_string_buffer(256)
    ref.ref_strref_message(*args, buf, 256)
    return buf.value.decode()


@pytest.mark.parametrize("case", ["table", "column", "page"])
def test_oracle_execute_strref_overflow_message(oracle, ref, case):
    """the oracle's execute() fails where the reference's copy_scan throws, with the message the
    reference's StrRef constructor throws (compiled from include/value_t.h)"""
    plan, bad = hjtest.strref_plan(oracle.codec, case)
    want = ref_message(ref, bad)
    assert want.startswith("StrRef " + case + " overflow")
    with pytest.raises(RuntimeError) as e:
        oracle.execute(plan)
    assert e.value.args[0].endswith(": " + want)


@pytest.mark.parametrize("case", ["page_null", "none"])
def test_oracle_execute_strref_within_limits(oracle, case):
    """no StrRef is made for NULL rows, and page 2^18 - 1 still fits"""
    plan, _ = hjtest.strref_plan(oracle.codec, case)
    rows = oracle.rows(oracle.execute(plan))
    assert len(rows) == hjtest.STRREF_PAGE_LIMIT + (1 if case == "page_null" else 0)
    assert sum(1 for r in rows if r == ("a",)) == (1 if case == "none" else 0)
