"""column_t, the reference's intermediate result column (include/column_t.h:16-82): the cases of
tests/columnt_tests.cpp:5-204 against the oracle's restatement, plus its direct mode over a dense
input column and the chunk merge (include/column_merge.h:7-31)."""
import ctypes as C

import numpy as np
import pytest

import hjtest

INT32, VARCHAR = int(hjtest.DataType.INT32), int(hjtest.DataType.VARCHAR)


class Col:
    def __init__(self, oracle, type):
        L = oracle.lib
        for name, res, args in (
                ("oracle_vcol_new", C.c_void_p, [C.c_int32]),
                ("oracle_vcol_append", None, [C.c_void_p, C.c_uint64]),
                ("oracle_vcol_get", C.c_uint64, [C.c_void_p, C.c_uint64]),
                ("oracle_vcol_rows", C.c_uint64, [C.c_void_p]),
                ("oracle_vcol_pages", C.c_uint64, [C.c_void_p]),
                ("oracle_vcol_type", C.c_int32, [C.c_void_p]),
                ("oracle_vcol_use_input", None, [C.c_void_p, C.c_void_p, C.c_uint64]),
                ("oracle_vcol_merge", None, [C.c_void_p, C.c_void_p]),
                ("oracle_vcol_free", None, [C.c_void_p])):
            f = getattr(L, name)
            f.restype, f.argtypes = res, args
        self.L, self.oracle = L, oracle
        self.h = L.oracle_vcol_new(type)
        self.keep = []

    def __del__(self):
        self.L.oracle_vcol_free(self.h)

    def append_int(self, v):
        self.L.oracle_vcol_append(self.h, self.L.oracle_value_int32(v))

    def append_null(self):
        self.L.oracle_vcol_append(self.h, 0)  # value_t() is NULL: all bits zero

    def append(self, value):
        self.L.oracle_vcol_append(self.h, value)

    def get(self, row):
        return self.L.oracle_vcol_get(self.h, row)

    def get_int(self, row):
        v = self.get(row)
        assert v >> 62 == 1, "is_int32"
        return C.c_int32(v & 0xFFFFFFFF).value

    @property
    def rows(self):
        return self.L.oracle_vcol_rows(self.h)

    @property
    def pages(self):
        return self.L.oracle_vcol_pages(self.h)

    @property
    def type(self):
        return self.L.oracle_vcol_type(self.h)


def test_basic_construction(oracle):
    # columnt_tests.cpp:5-18
    for t in (INT32, VARCHAR):
        c = Col(oracle, t)
        assert c.type == t and c.rows == 0 and c.pages == 0


def test_append_single_value(oracle):
    # columnt_tests.cpp:20-34
    c = Col(oracle, INT32)
    c.append_int(42)
    assert c.rows == 1 and c.pages == 1
    assert c.get_int(0) == 42


@pytest.mark.parametrize("n", [10, 100])
def test_append_multiple_values(oracle, n):
    # columnt_tests.cpp:36-71
    c = Col(oracle, INT32)
    for i in range(n):
        c.append_int(i * 10 if n == 10 else i)
    assert c.rows == n and c.pages == 1
    for i in range(n):
        assert c.get_int(i) == (i * 10 if n == 10 else i)


@pytest.mark.parametrize("n,pages", [(1024, 1), (1025, 2), (3000, 3)])
def test_page_allocation(oracle, n, pages):
    # columnt_tests.cpp:73-122: 1024 values per page
    c = Col(oracle, INT32)
    for i in range(n):
        c.append_int(i)
    assert c.rows == n and c.pages == pages
    for r in (0, 1023, min(1024, n - 1), n // 2, n - 1):
        assert c.get_int(r) == r


def test_null_values(oracle):
    # columnt_tests.cpp:124-146
    c = Col(oracle, INT32)
    c.append_int(99)
    c.append_null()
    c.append_int(99)
    assert c.rows == 3
    assert c.get_int(0) == 99 and c.get(1) == 0 and c.get_int(2) == 99


def test_strref_values(oracle):
    # columnt_tests.cpp:148-177: StrRef(false, 1, 2, 3, 4) and StrRef(true, 5, 6, 7, 8)
    c = Col(oracle, VARCHAR)
    refs = []
    for args in ((0, 1, 2, 3, 4), (1, 5, 6, 7, 8)):
        out = C.c_uint64()
        assert oracle.lib.oracle_strref_encode(*args, C.byref(out)) == 0
        refs.append(out.value)
        c.append(oracle.lib.oracle_value_strref(out.value))
    assert c.rows == 2
    r1, r2 = c.get(0), c.get(1)
    assert r1 >> 62 == 2 and r2 >> 62 == 2  # is_strref
    assert (r1 >> 51) & 1023 == 1 and (r1 >> 41) & 1023 == 2  # table, column
    assert (r2 >> 61) & 1 == 1 and (r2 >> 51) & 1023 == 5  # is_long, table


def test_edge_cases(oracle):
    # columnt_tests.cpp:179-204
    c = Col(oracle, INT32)
    assert c.rows == 0 and c.pages == 0
    for _ in range(1024):
        c.append_int(0)
    assert c.pages == 1
    c.append_int(999)
    assert c.pages == 2 and c.rows == 1025 and c.get_int(1024) == 999


def test_direct_mode_reads_dense_input_pages(oracle):
    """use_input_col (column_t.h:76-80): rows of a dense INT32 column read in place at page
    row / 1984, slot row % 1984, which is where the page writer puts them (1984 values fill a
    page without NULLs: 4 + 4 * 1984 + 248 = 8188 bytes, one more value does not fit)"""
    v = np.random.default_rng(3).integers(-2**31, 2**31, 5000).astype(np.int32)
    col = oracle.codec.encode_int32(v)
    assert col.num_pages == 3  # 1984 + 1984 + 1032
    keep = []
    hc = hjtest._plan.column_to_c(col, keep)
    c = Col(oracle, INT32)
    c.L.oracle_vcol_use_input(c.h, C.addressof(hc), len(v))
    assert c.rows == 5000
    for r in (0, 1983, 1984, 3967, 3968, 4999):
        assert c.get_int(r) == v[r]


def test_merge_moves_pages_into_an_empty_column_then_appends(oracle):
    """merge_column_pages (column_merge.h:7-31): the first chunk's pages move, later chunks
    append row by row after the partly filled last page"""
    dst = Col(oracle, INT32)
    a, b = Col(oracle, INT32), Col(oracle, INT32)
    for i in range(1500):
        a.append_int(i)
    for i in range(1500, 2600):
        b.append_int(i)
    dst.L.oracle_vcol_merge(dst.h, a.h)
    assert dst.rows == 1500 and dst.pages == 2 and a.rows == 0 and a.pages == 0
    dst.L.oracle_vcol_merge(dst.h, b.h)
    assert dst.rows == 2600 and dst.pages == 3  # 1024 + 1024 + 552 values
    for r in (0, 1023, 1499, 1500, 2047, 2048, 2599):
        assert dst.get_int(r) == r
