"""Python mirror of the reference's plan/table API (include/plan.h:73-210, include/table.h:39-125).

Plan, PlanNode, ScanNode, JoinNode, Column, ColumnarTable and Table keep the reference's names and
argument meaning so plans and tests read like the reference's own (tests/unit_tests.cpp). Pages are
numpy arrays of shape (num_pages, 8192), byte-identical to the reference's Page objects.
"""
import ctypes as C
from enum import IntEnum

import numpy as np

from . import abi

PAGE_SIZE = abi.PAGE_SIZE


class DataType(IntEnum):
    """include/attribute.h:8-13"""
    INT32 = 0
    INT64 = 1
    FP64 = 2
    VARCHAR = 3


class Column:
    """include/plan.h:106-161: a typed sequence of 8192-byte pages."""

    def __init__(self, type, pages=None, page_index=None):
        self.type = DataType(type)
        if pages is None:
            pages = np.zeros((0, PAGE_SIZE), dtype=np.uint8)
        self.pages = np.ascontiguousarray(pages, dtype=np.uint8).reshape(-1, PAGE_SIZE)
        # Column::pages holds page pointers, and one page may stand at several positions:
        # page i of the column is pages[page_index[i]] (None: pages in order)
        self.page_index = None if page_index is None else np.asarray(page_index, dtype=np.int64)

    @property
    def num_pages(self):
        return self.pages.shape[0] if self.page_index is None else len(self.page_index)


class ColumnarTable:
    """include/plan.h:163-166"""

    def __init__(self, num_rows=0, columns=None):
        self.num_rows = int(num_rows)
        self.columns = list(columns or [])


class ScanNode:
    """include/plan.h:78-80, plus engine-side range filters (lo <= col <= hi, NULL fails) and a
    predicate program: the reference's Comparison / LogicalOperation tree (include/statement.h)
    in postfix order, entries ("eq"|"neq"|"lt"|"gt"|"leq"|"geq", col, literal),
    ("like"|"not_like", col, pattern), ("is_null"|"is_not_null", col), ("and",), ("or",),
    ("not",)."""

    def __init__(self, base_table_id, filters=(), preds=()):
        self.base_table_id = int(base_table_id)
        self.filters = [(int(c), int(lo), int(hi)) for c, lo, hi in filters]
        self.preds = [tuple(p) for p in preds]


class JoinNode:
    """include/plan.h:82-88"""

    def __init__(self, build_left, left, right, left_attr, right_attr):
        self.build_left = bool(build_left)
        self.left = int(left)
        self.right = int(right)
        self.left_attr = int(left_attr)
        self.right_attr = int(right_attr)


class PlanNode:
    """include/plan.h:90-98"""

    def __init__(self, data, output_attrs):
        self.data = data
        self.output_attrs = [(int(i), DataType(t)) for i, t in output_attrs]


class Plan:
    """include/plan.h:173-210"""

    def __init__(self):
        self.nodes = []
        self.inputs = []
        self.root = 0

    def new_join_node(self, build_left, left, right, left_attr, right_attr, output_attrs):
        self.nodes.append(PlanNode(JoinNode(build_left, left, right, left_attr, right_attr),
                                   output_attrs))
        return len(self.nodes) - 1

    def new_scan_node(self, base_table_id, output_attrs, filters=(), preds=()):
        self.nodes.append(PlanNode(ScanNode(base_table_id, filters, preds), output_attrs))
        return len(self.nodes) - 1

    def new_input(self, table):
        self.inputs.append(table)
        return len(self.inputs) - 1

    # ---- marshalling to the C ABI (include/hjoin.h)
    def to_c(self, with_inputs=True):
        """Returns (hj_plan, keepalive). The keepalive list must outlive every use of hj_plan."""
        keep = []
        nodes = (abi.hj_node * max(1, len(self.nodes)))()
        for i, nd in enumerate(self.nodes):
            n = nodes[i]
            attrs = (abi.hj_attr * max(1, len(nd.output_attrs)))()
            for j, (idx, t) in enumerate(nd.output_attrs):
                attrs[j].index = idx
                attrs[j].type = int(t)
            keep.append(attrs)
            n.num_output_attrs = len(nd.output_attrs)
            n.output_attrs = attrs
            if isinstance(nd.data, ScanNode):
                n.kind = abi.NODE_SCAN
                n.base_table_id = nd.data.base_table_id
                fl = (abi.hj_filter * max(1, len(nd.data.filters)))()
                for j, (c, lo, hi) in enumerate(nd.data.filters):
                    fl[j].column, fl[j].lo, fl[j].hi = c, lo, hi
                keep.append(fl)
                n.num_filters = len(nd.data.filters)
                n.filters = fl
                pr = (abi.hj_pred * max(1, len(nd.data.preds)))()
                for j, e in enumerate(nd.data.preds):
                    pr[j].op = abi.PRED_OPS[e[0]]
                    if len(e) > 1:
                        pr[j].column = int(e[1])
                    if len(e) > 2:
                        lit = e[2]
                        if isinstance(lit, (str, bytes)):
                            b = lit.encode() if isinstance(lit, str) else bytes(lit)
                            buf = C.create_string_buffer(b, len(b) + 1)
                            keep.append(buf)
                            pr[j].sval = C.cast(buf, C.c_char_p)
                            pr[j].slen = len(b)
                        else:
                            pr[j].ival = int(lit)
                keep.append(pr)
                n.num_preds = len(nd.data.preds)
                n.preds = pr
            else:
                d = nd.data
                n.kind = abi.NODE_JOIN
                n.build_left = int(d.build_left)
                n.left, n.right = d.left, d.right
                n.left_attr, n.right_attr = d.left_attr, d.right_attr
        p = abi.hj_plan()
        p.num_nodes = len(self.nodes)
        p.nodes = nodes
        p.root = self.root
        keep.append(nodes)
        p.num_inputs = len(self.inputs)
        if with_inputs:
            tabs, tkeep = tables_to_c(self.inputs)
            keep.extend(tkeep)
            p.inputs = tabs
        return p, keep


def column_to_c(col, keep):
    ptrs = (C.c_void_p * max(1, col.num_pages))()
    base = col.pages.ctypes.data
    if col.page_index is None:
        for i in range(col.num_pages):
            ptrs[i] = base + i * PAGE_SIZE
    elif col.num_pages:
        addr = (base + col.page_index * PAGE_SIZE).astype(np.uint64)
        C.memmove(ptrs, addr.ctypes.data, 8 * col.num_pages)
        keep.append(addr)
    keep.append(ptrs)
    keep.append(col.pages)
    c = abi.hj_column()
    c.type = int(col.type)
    c.num_pages = col.num_pages
    c.pages = ptrs
    return c


def table_to_c(t, keep):
    cols = (abi.hj_column * max(1, len(t.columns)))()
    for j, col in enumerate(t.columns):
        cols[j] = column_to_c(col, keep)
    keep.append(cols)
    ht = abi.hj_table()
    ht.num_rows = t.num_rows
    ht.num_columns = len(t.columns)
    ht.columns = cols
    return ht


def tables_to_c(tables):
    keep = []
    arr = (abi.hj_table * max(1, len(tables)))()
    for i, t in enumerate(tables):
        arr[i] = table_to_c(t, keep)
    keep.append(arr)
    return arr, keep


# ------------------------------------------------------------------------------ .tbl cache files
class CacheError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code
        self.msg = msg


def write_cache(table, path, lib=None):
    """DumpTable::dump (include/table.h:200-211): the header page (TableMeta), then every
    column's pages, to `path` (hj_cache_write)"""
    lib = lib or abi.load()
    keep = []
    ht = table_to_c(table, keep)
    rc = lib.hj_cache_write(C.byref(ht), str(path).encode())
    if rc:
        raise CacheError(rc, lib.hj_global_error().decode())


def read_cache(path, lib=None):
    """Table::from_cache (src/build_table.cpp:147-203): the file's ColumnarTable, its pages
    copied out of the mapping (hj_cache_open)"""
    lib = lib or abi.load()
    h = C.c_void_p()
    rc = lib.hj_cache_open(str(path).encode(), C.byref(h))
    if rc:
        raise CacheError(rc, lib.hj_global_error().decode())
    try:
        t = lib.hj_cache_table(h).contents
        cols = []
        for j in range(t.num_columns):
            c = t.columns[j]
            pages = np.empty((c.num_pages, PAGE_SIZE), dtype=np.uint8)
            if c.num_pages:  # a column's pages are contiguous in the file
                C.memmove(pages.ctypes.data, c.pages[0], c.num_pages * PAGE_SIZE)
            cols.append(Column(c.type, pages))
        return ColumnarTable(t.num_rows, cols)
    finally:
        lib.hj_cache_close(h)


# ------------------------------------------------------------------------------ page codec
class Codec:
    """Page writer/reader behind a C library exposing <prefix>encode_int32 etc. with the
    signatures of include/hjoin.h (the engine's hj_ functions by default)."""

    def __init__(self, lib=None, prefix="hj_"):
        self.lib = lib if lib is not None else abi.load()
        self.p = prefix
        L = self.lib
        self._enc_i = getattr(L, prefix + "encode_int32")
        self._enc_s = getattr(L, prefix + "encode_varchar")
        self._free = getattr(L, prefix + "out_column_free")
        self._rows = getattr(L, prefix + ("column_num_rows" if prefix == "hj_" else "column_rows"))
        self._chars = getattr(L, prefix + ("column_num_chars" if prefix == "hj_" else "column_chars"))
        self._dec_i = getattr(L, prefix + "decode_int32")
        self._dec_s = getattr(L, prefix + "decode_varchar")
        for f, args, res in (
                (self._enc_i, [C.c_void_p, C.c_void_p, C.c_uint64, C.POINTER(abi.hj_out_column)], C.c_int),
                (self._enc_s, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64,
                               C.POINTER(abi.hj_out_column)], C.c_int),
                (self._free, [C.POINTER(abi.hj_out_column)], None),
                (self._rows, [C.POINTER(abi.hj_column)], C.c_uint64),
                (self._chars, [C.POINTER(abi.hj_column)], C.c_uint64),
                (self._dec_i, [C.POINTER(abi.hj_column), C.c_void_p, C.c_void_p], C.c_int),
                (self._dec_s, [C.POINTER(abi.hj_column), C.c_void_p, C.c_void_p, C.c_void_p],
                 C.c_int)):
            f.argtypes = args
            f.restype = res

    def _take(self, oc, type):
        n = oc.num_pages
        pages = np.empty((n, PAGE_SIZE), dtype=np.uint8)
        if n:
            C.memmove(pages.ctypes.data, oc.pages, n * PAGE_SIZE)
        self._free(C.byref(oc))
        return Column(type, pages)

    def encode_int32(self, values, validity=None):
        v = np.ascontiguousarray(values, dtype=np.int32)
        ok = None if validity is None else np.ascontiguousarray(validity, dtype=np.uint8)
        oc = abi.hj_out_column()
        rc = self._enc_i(v.ctypes.data if len(v) else None,
                         ok.ctypes.data if ok is not None and len(ok) else None, len(v),
                         C.byref(oc))
        if rc:
            raise RuntimeError(f"encode_int32 failed ({rc})")
        return self._take(oc, DataType.INT32)

    def encode_varchar(self, offsets, chars, validity=None):
        off = np.ascontiguousarray(offsets, dtype=np.uint64)
        ch = np.frombuffer(chars, dtype=np.uint8) if isinstance(chars, (bytes, bytearray)) \
            else np.ascontiguousarray(chars, dtype=np.uint8)
        if len(ch) == 0:
            ch = np.zeros(1, dtype=np.uint8)
        n = len(off) - 1
        ok = None if validity is None else np.ascontiguousarray(validity, dtype=np.uint8)
        oc = abi.hj_out_column()
        rc = self._enc_s(off.ctypes.data, ch.ctypes.data,
                         ok.ctypes.data if ok is not None and n else None, n, C.byref(oc))
        if rc:
            raise RuntimeError(f"encode_varchar failed ({rc})")
        return self._take(oc, DataType.VARCHAR)

    def encode_strings(self, strings):
        """list of str/bytes/None -> VARCHAR column"""
        bs = [b"" if s is None else (s.encode() if isinstance(s, str) else bytes(s))
              for s in strings]
        off = np.zeros(len(bs) + 1, dtype=np.uint64)
        if bs:
            off[1:] = np.cumsum([len(b) for b in bs])
        ok = np.array([s is not None for s in strings], dtype=np.uint8)
        return self.encode_varchar(off, b"".join(bs), ok)

    def decode(self, col):
        keep = []
        hc = column_to_c(col, keep)
        n = int(self._rows(C.byref(hc)))
        ok = np.zeros(max(n, 1), dtype=np.uint8)
        if col.type == DataType.INT32:
            v = np.zeros(max(n, 1), dtype=np.int32)
            rc = self._dec_i(C.byref(hc), v.ctypes.data, ok.ctypes.data)
            if rc:
                raise RuntimeError(f"decode_int32 failed ({rc})")
            return v[:n], ok[:n].astype(bool)
        if col.type == DataType.VARCHAR:
            nc = int(self._chars(C.byref(hc)))
            off = np.zeros(n + 1, dtype=np.uint64)
            ch = np.zeros(max(nc, 1), dtype=np.uint8)
            rc = self._dec_s(C.byref(hc), off.ctypes.data, ch.ctypes.data, ok.ctypes.data)
            if rc:
                raise RuntimeError(f"decode_varchar failed ({rc})")
            return (off, ch[:nc].tobytes()), ok[:n].astype(bool)
        raise TypeError(f"cannot decode {col.type!r}")

    def column_values(self, col):
        """decoded column as a list of python values (int / str / None)"""
        vals, ok = self.decode(col)
        if col.type == DataType.INT32:
            return [int(x) if o else None for x, o in zip(vals.tolist(), ok.tolist())]
        off, ch = vals
        return [ch[off[i]:off[i + 1]].decode("latin-1") if ok[i] else None
                for i in range(len(ok))]


class Table:
    """include/table.h:39-125: rows of Data (int / str / None) plus column types."""

    def __init__(self, rows, types):
        self.rows = [list(r) for r in rows]
        self.types = [DataType(t) for t in types]

    def to_columnar(self, codec=None):
        """src/build_table.cpp:657-882"""
        codec = codec or Codec()
        cols = []
        for j, t in enumerate(self.types):
            vals = [r[j] for r in self.rows]
            if t == DataType.INT32:
                v = np.array([0 if x is None else x for x in vals], dtype=np.int64).astype(np.int32)
                ok = np.array([x is not None for x in vals], dtype=np.uint8)
                cols.append(codec.encode_int32(v, ok))
            elif t == DataType.VARCHAR:
                cols.append(codec.encode_strings(vals))
            else:
                raise TypeError("only INT32 and VARCHAR columns are produced here")
        return ColumnarTable(len(self.rows), cols)

    @staticmethod
    def from_columnar(ct, codec=None):
        """src/build_table.cpp:513-637 -> Table"""
        codec = codec or Codec()
        cols = [codec.column_values(c) for c in ct.columns]
        rows = [list(r) for r in zip(*cols)] if cols else []
        if cols and len(rows) != ct.num_rows:
            raise ValueError("column row counts differ from num_rows")
        return Table(rows, [c.type for c in ct.columns])


def sort_key(row):
    """total order over rows with NULLs (None sorts first, like std::monostate compares
    after the other alternatives is irrelevant for multiset equality)"""
    return tuple((0, 0) if v is None else (1, v) if isinstance(v, int) else (2, v) for v in row)


def sorted_rows(table):
    return sorted((tuple(r) for r in table.rows), key=sort_key)


def result_from_c(res, free_fn):
    """hj_result -> ColumnarTable (copies pages), then frees the C result with free_fn"""
    cols = []
    for j in range(res.num_columns):
        oc = res.columns[j]
        n = oc.num_pages
        pages = np.empty((n, PAGE_SIZE), dtype=np.uint8)
        if n:
            C.memmove(pages.ctypes.data, oc.pages, n * PAGE_SIZE)
        cols.append(Column(oc.type, pages))
    ct = ColumnarTable(res.num_rows, cols)
    free_fn(C.byref(res))
    return ct
