""".tbl cache files (Table::from_cache, src/build_table.cpp:147-203; DumpTable::dump,
include/table.h:128-211): the engine's writer and reader against the oracle's restatement and
against the header layout that TableMeta's C definition fixes (include/table.h:30-37). Host only;
the resident upload of a cache file runs in test_gpu_parity."""
import struct

import numpy as np
import pytest

import hjtest

PAGE = 8192
I32, STR = hjtest.DataType.INT32, hjtest.DataType.VARCHAR


def _table(oracle, seed, ncols=3, rows=5000, long_every=0):
    rng = np.random.default_rng(seed)
    t = hjtest.random_table(rng, rows, max(0, ncols - 1), 1000, null_p=0.2, with_str=ncols > 0,
                            long_every=long_every)
    return t.to_columnar(oracle.codec)


def _same(a, b):
    assert a.num_rows == b.num_rows
    assert len(a.columns) == len(b.columns)
    for x, y in zip(a.columns, b.columns):
        assert int(x.type) == int(y.type)
        assert x.num_pages == y.num_pages
        assert np.array_equal(x.pages, y.pages)


def test_header_layout(hjo, oracle, tmp_path):
    # TableMeta {u64 num_rows; u64 num_cols; DataType types[16]; u64 num_pages[16]} with an
    # int-sized enum: fields at 0, 8, 16 and 80, 208 bytes, then zeros to the end of the page
    t = _table(oracle, 1, ncols=2, rows=3000)
    path = tmp_path / "t.tbl"
    hjo.write_cache(t, path)
    raw = path.read_bytes()
    npages = [c.num_pages for c in t.columns]
    assert len(raw) == PAGE * (1 + sum(npages))
    assert struct.unpack_from("<QQ", raw, 0) == (3000, 2)
    assert list(struct.unpack_from("<16i", raw, 16)) == [int(I32), int(STR)] + [0] * 14
    assert list(struct.unpack_from("<16Q", raw, 80)) == npages + [0] * 14
    assert raw[208:PAGE] == bytes(PAGE - 208)
    body = np.frombuffer(raw, np.uint8)[PAGE:].reshape(-1, PAGE)
    assert np.array_equal(body[:npages[0]], t.columns[0].pages)
    assert np.array_equal(body[npages[0]:], t.columns[1].pages)


@pytest.mark.parametrize("case", ["plain", "long_strings", "no_columns", "empty_rows",
                                  "sixteen_columns"])
def test_writer_and_reader_match_the_oracle(hjo, oracle, tmp_path, case):
    if case == "plain":
        t = _table(oracle, 2)
    elif case == "long_strings":
        t = _table(oracle, 3, long_every=50, rows=600)
    elif case == "no_columns":
        t = hjtest.ColumnarTable(0, [])
    elif case == "empty_rows":
        t = hjtest.Table([], [I32, STR]).to_columnar(oracle.codec)
    else:
        rng = np.random.default_rng(4)
        t = hjtest.random_table(rng, 700, 15, 50, null_p=0.3).to_columnar(oracle.codec)
        assert len(t.columns) == 16
    mine, theirs = tmp_path / "mine.tbl", tmp_path / "oracle.tbl"
    hjo.write_cache(t, mine)
    oracle.cache_write(t, theirs)
    assert mine.read_bytes() == theirs.read_bytes()
    _same(hjo.read_cache(theirs), t)   # the engine's reader on the oracle's file
    _same(oracle.cache_read(mine), t)  # and the other way round


def test_rows_survive_the_round_trip(hjo, oracle, tmp_path):
    t = _table(oracle, 5, ncols=3, rows=4000, long_every=700)
    path = tmp_path / "t.tbl"
    hjo.write_cache(t, path)
    back = hjo.read_cache(path)
    assert oracle.rows(back) == oracle.rows(t)


def test_errors_carry_the_reference_messages(hjo, oracle, tmp_path):
    missing = tmp_path / "missing.tbl"
    with pytest.raises(hjo.CacheError, match=f"Failed to open file: {missing}"):
        hjo.read_cache(missing)
    empty = tmp_path / "empty.tbl"
    empty.write_bytes(b"")
    with pytest.raises(hjo.CacheError, match=f"File is empty: {empty}"):
        hjo.read_cache(empty)


def test_malformed_files_are_refused(hjo, oracle, tmp_path):
    t = _table(oracle, 6, ncols=2, rows=3000)
    good = tmp_path / "good.tbl"
    hjo.write_cache(t, good)
    raw = bytearray(good.read_bytes())
    cases = {
        "truncated": raw[:-100],
        "extra_page": raw + bytes(PAGE),
        "short_header": raw[:PAGE - 1],
        "seventeen_columns": raw[:8] + struct.pack("<Q", 17) + raw[16:],
        "unknown_type": raw[:16] + struct.pack("<i", 7) + raw[20:],
    }
    for name, data in cases.items():
        p = tmp_path / f"{name}.tbl"
        p.write_bytes(bytes(data))
        with pytest.raises(hjo.CacheError):
            hjo.read_cache(p)
        if name != "unknown_type":  # the oracle, like the reference, does not check types
            assert oracle.cache_read(p) is None, name


def test_more_than_sixteen_columns_cannot_be_written(hjo, oracle, tmp_path):
    rng = np.random.default_rng(7)
    t = hjtest.random_table(rng, 10, 16, 5, with_str=True).to_columnar(oracle.codec)
    assert len(t.columns) == 17
    with pytest.raises(hjo.CacheError):
        hjo.write_cache(t, tmp_path / "x.tbl")
