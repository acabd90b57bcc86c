"""CPU ORACLE -- test infrastructure only.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may import this package, and
only as the checker (or the timed CPU baseline). The product path (ghive_amd) never imports it.

numpy front end over liboracle.so (oracle/ghive_oracle.c), a C restatement of Hive 3.1's
vectorized operators and of the reference GPU library's operator contracts; see the header of
ghive_oracle.h for the reference file:line each function follows.
"""
import ctypes as C
import os
import subprocess

import numpy as np

from ghive_amd import abi

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_NP_OF_TYPE = {abi.LONG: np.int64, abi.DOUBLE: np.float64, abi.INT: np.int32,
               abi.FLOAT: np.float32}
_TYPE_OF_NP = {np.dtype(np.int64): abi.LONG, np.dtype(np.float64): abi.DOUBLE,
               np.dtype(np.int32): abi.INT, np.dtype(np.float32): abi.FLOAT}


def build(force=False):
    """Compile liboracle.so with the committed Makefile (gcc)."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib():
    global _lib
    if _lib is None:
        build()
        L = C.CDLL(_LIB_PATH)
        P, u64, u32, i64 = C.c_void_p, C.c_uint64, C.c_uint32, C.c_int64
        L.orc_philox4x32_10.argtypes = [P, P, P]
        for name, extra in (("orc_gen_uniform_i32", [C.c_int32, C.c_int32]),
                            ("orc_gen_uniform_i64", [C.c_int64, C.c_int64]),
                            ("orc_gen_uniform_f32", []), ("orc_gen_normal_f64", [])):
            fn = getattr(L, name)
            fn.argtypes = [u64, u32, u64, u64] + extra + [P]
            fn.restype = None
        L.orc_key62.argtypes = [u64, u64]
        L.orc_key62.restype = u64
        L.orc_gen_zipf_keys.argtypes = [u64, u32, u64, u64, u64, C.c_double, C.c_double, P, P]
        L.orc_gen_zipf_keys.restype = None
        L.orc_hive_long_hash.argtypes = [i64]
        L.orc_hive_long_hash.restype = C.c_int32
        L.orc_filter.argtypes = [P, P, u32, u32, P]
        L.orc_filter.restype = i64
        L.orc_mj_build.argtypes = [P, P, u64, P, C.c_float]
        L.orc_mj_build.restype = P
        L.orc_mj_free.argtypes = [P]
        L.orc_mj_stats.argtypes = [P, P, P, P, P, P]
        L.orc_mj_probe.argtypes = [P, P, P, u64, P, P, u64, P]
        L.orc_mj_probe.restype = i64
        L.orc_join.argtypes = [P, u32, P, P, u32, P, u64, P]
        L.orc_join.restype = i64
        L.orc_groupby.argtypes = [P, P, u64, P, u32, P, u32, C.c_int, P, P, P, P, u64]
        L.orc_groupby.restype = i64
        L.orc_exact_sum.argtypes = [P, u64]
        L.orc_exact_sum.restype = C.c_double
        L.orc_sort.argtypes = [P, P, P, u32, P]
        L.orc_sort.restype = C.c_int
        L.orc_sort_runs.argtypes = [P, P, P, u32, C.c_int, P, P, P]
        L.orc_sort_runs.restype = C.c_int
        L.orc_fragment_run.argtypes = [P, C.c_int, P, P, P, u64, P, P]
        L.orc_fragment_run.restype = i64
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


class HostTable:
    """Host columns (numpy) described as a ghive_table; keeps the arrays alive."""

    def __init__(self, columns):
        self.arrays = []
        n = len(columns)
        self.cols = (abi.Column * max(n, 1))()
        rows = None
        for i, col in enumerate(columns):
            if isinstance(col, tuple) and col[0] == "string":
                blob, idx = col[1], np.ascontiguousarray(col[2], dtype=np.int32)
                self.arrays += [blob, idx]
                self.cols[i].type = abi.STRING
                self.cols[i].data = C.cast(C.c_char_p(blob), C.c_void_p)
                self.cols[i].aux = idx.ctypes.data
                self.cols[i].row_num = len(idx) // 2
                self.cols[i].char_size = len(blob)
                r = len(idx) // 2
            else:
                a = np.ascontiguousarray(col)
                self.arrays.append(a)
                self.cols[i].type = _TYPE_OF_NP[a.dtype]
                self.cols[i].data = a.ctypes.data
                self.cols[i].row_num = len(a)
                r = len(a)
            rows = r if rows is None else rows
            assert rows == r, "ragged table"
        self.table = abi.Table(self.cols, n, 0, rows or 0)

    def ref(self):
        return C.byref(self.table)


def philox(ctr, key):
    c = np.ascontiguousarray(ctr, dtype=np.uint32)
    k = np.ascontiguousarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def gen_uniform_i32(seed, col, row_begin, n, lo, hi):
    out = np.empty(n, dtype=np.int32)
    lib().orc_gen_uniform_i32(seed, col, row_begin, n, lo, hi, _ptr(out))
    return out


def gen_uniform_i64(seed, col, row_begin, n, lo, hi):
    out = np.empty(n, dtype=np.int64)
    lib().orc_gen_uniform_i64(seed, col, row_begin, n, lo, hi, _ptr(out))
    return out


def gen_uniform_f32(seed, col, row_begin, n):
    out = np.empty(n, dtype=np.float32)
    lib().orc_gen_uniform_f32(seed, col, row_begin, n, _ptr(out))
    return out


def gen_normal_f64(seed, col, row_begin, n):
    out = np.empty(n, dtype=np.float64)
    lib().orc_gen_normal_f64(seed, col, row_begin, n, _ptr(out))
    return out


def key62(seed, rank):
    return lib().orc_key62(seed, rank)


def gen_zipf_keys(seed, col, row_begin, n, nranks, theta, miss_frac):
    """(keys int64, ranks uint64) of the skewed probe side."""
    keys = np.empty(n, dtype=np.int64)
    ranks = np.empty(n, dtype=np.uint64)
    lib().orc_gen_zipf_keys(seed, col, row_begin, n, nranks, theta, miss_frac, _ptr(keys), _ptr(ranks))
    return keys, ranks


def hive_long_hash(key):
    return lib().orc_hive_long_hash(int(key))


def filter_rows(columns, nodes, root):
    """Ids of rows passing the predicate tree (ascending)."""
    t = HostTable(columns)
    arr = abi.pred_array(nodes)
    out = np.empty(max(t.table.row_num, 1), dtype=np.uint32)
    n = lib().orc_filter(t.ref(), arr, len(nodes), root, _ptr(out))
    if n < 0:
        raise ValueError("oracle: unsupported predicate")
    return out[:n].copy()


_NULL_OF_NP = {np.dtype(np.int32): 2**31 - 1, np.dtype(np.int64): 2**63 - 1}


def select_arith(a, b, op):
    """Projection arithmetic a op b (CalculationSelectPredicate::process,
    gpu_impl/src/Operator/SelectPredicate.cu:84-409; functors SelectPredicate.hpp:193-234) with
    Hive's typing: '/' and any floating operand compute in float64; int32 op int32 stays int32
    and wraps like a Java int; other integer pairs int64 (wrapping like a Java long). NULL (the
    type's sentinel, NaN for floats) in either operand gives NULL; x / 0 gives NULL. A 1-element
    operand broadcasts."""
    a, b = np.asarray(a), np.asarray(b)
    n = max(len(a), len(b)) if min(len(a), len(b)) == 1 else len(a)
    ia = np.zeros(n, dtype=np.int64) if len(a) == 1 else np.arange(n)
    ib = np.zeros(n, dtype=np.int64) if len(b) == 1 else np.arange(n)
    a, b = a[ia], b[ib]

    def nulls(x):
        return np.isnan(x) if x.dtype.kind == "f" else x == _NULL_OF_NP[x.dtype]
    null = nulls(a) | nulls(b)
    if op == "/" or a.dtype.kind == "f" or b.dtype.kind == "f":
        x, y = a.astype(np.float64), b.astype(np.float64)
        with np.errstate(all="ignore"):
            r = {"+": np.add, "-": np.subtract, "*": np.multiply, "/": np.divide}[op](x, y)
        if op == "/":
            null |= y == 0.0
        r[null] = np.nan
        return r
    dt = np.int32 if a.dtype == b.dtype == np.int32 else np.int64
    x, y = a.astype(np.int64).astype(np.uint64), b.astype(np.int64).astype(np.uint64)
    r = {"+": np.add, "-": np.subtract, "*": np.multiply}[op](x, y)  # modulo 2^64
    r = (r & np.uint64(0xFFFFFFFF)).astype(np.uint32).view(np.int32) if dt == np.int32 else r.view(np.int64)
    r = r.copy()
    r[null] = _NULL_OF_NP[np.dtype(dt)]
    return r


_JAVA_ASCII_SPACE = set(b" \t\n\x0b\x0c\r\x1c\x1d\x1e\x1f")


def cast_string_to_int(values, bits=32):
    """CAST(string AS INT / BIGINT) as Hive 3.1: UDFToInteger.evaluate(Text)
    (ql/src/java/org/apache/hadoop/hive/ql/udf/UDFToInteger.java:169-186) = LazyUtils.isNumberMaybe
    (serde/.../lazy/LazyUtils.java:92-110) then LazyInteger.parseInt(bytes, 0, len, 10, trim=true)
    (serde/.../lazy/LazyInteger.java). values: bytes or None; returns int or None (NULL)."""
    lo, hi = -(1 << (bits - 1)), (1 << (bits - 1)) - 1
    out = []
    for v in values:
        out.append(None if v is None else _parse_java_int(bytes(v), lo, hi))
    return out


def _parse_java_int(b, lo, hi):
    n = len(b)
    isdig = lambda c: 48 <= c <= 57  # noqa: E731
    if n == 0 or (n == 1 and not isdig(b[0])) or (n == 2 and not (isdig(b[0]) or isdig(b[1]))) or \
            (n == 4 and b[0] in b"Nn"):
        return None
    s, e = 0, n
    while s < e and b[s] in _JAVA_ASCII_SPACE:
        s += 1
    while e > s and b[e - 1] in _JAVA_ASCII_SPACE:
        e -= 1
    if s == e:
        return None
    neg = b[s] == 0x2D
    i = s
    if neg or b[s] == 0x2B:
        i += 1
        if e - s == 1:
            return None
    whole = b[i:e].split(b".", 1)
    if not whole[0].isdigit() and whole[0] != b"":
        return None
    if len(whole) == 2 and whole[1] and not whole[1].isdigit():
        return None
    if len(whole) == 1 and whole[0] == b"":
        return None
    v = int(whole[0]) if whole[0] else 0
    v = -v if neg else v
    return v if lo <= v <= hi else None


def substr(values, pos, length):
    """SUBSTR(str, pos, len) as Hive 3.1's UDFSubstr.evaluate(Text, IntWritable, IntWritable)
    (ql/src/java/org/apache/hadoop/hive/ql/udf/UDFSubstr.java:61-105): Java String positions (UTF-16
    units), makeIndex's rules. values: bytes or None; returns bytes or None."""
    out = []
    for v in values:
        if v is None:
            out.append(None)
            continue
        u = bytes(v).decode("utf-8", errors="replace").encode("utf-16-le", errors="surrogatepass")
        L = len(u) // 2
        if length <= 0 or abs(pos) > L:
            out.append(b"")
            continue
        start = pos - 1 if pos > 0 else (L + pos if pos < 0 else 0)
        end = L if L - start < length else start + length
        out.append(u[2 * start: 2 * end].decode("utf-16-le", errors="surrogatepass").encode("utf-8"))
    return out


def mapjoin_inner(build_keys, probe_keys, build_sel=None, probe_sel=None, payload=None,
                  load_factor=0.75):
    """Hive map-join (inner) restated: (probe row ids, payloads) in Hive emission order."""
    bt = HostTable([build_keys] + ([payload] if payload is not None else []))
    pt = HostTable([probe_keys])
    bsel = None if build_sel is None else np.ascontiguousarray(build_sel, dtype=np.uint32)
    psel = None if probe_sel is None else np.ascontiguousarray(probe_sel, dtype=np.uint32)
    nb = len(bsel) if bsel is not None else len(build_keys)
    npr = len(psel) if psel is not None else len(probe_keys)
    L = lib()
    pcol = C.byref(bt.cols[1]) if payload is not None else None
    h = L.orc_mj_build(C.byref(bt.cols[0]), _ptr(bsel), nb, pcol, load_factor)
    if not h:
        raise ValueError("oracle: unsupported key type")
    try:
        need = C.c_uint64(0)
        cap = max(npr, 1)
        while True:
            rows = np.empty(cap, dtype=np.uint32)
            pay = np.empty(cap, dtype=np.int64)
            n = L.orc_mj_probe(h, C.byref(pt.cols[0]), _ptr(psel), npr, _ptr(rows), _ptr(pay),
                               cap, C.byref(need))
            if n >= 0:
                return rows[:n].copy(), pay[:n].copy()
            cap = need.value
    finally:
        L.orc_mj_free(h)


class MapJoinTable:
    """A built map-join hash table that several threads probe at once (the probe only reads the
    table, and ctypes releases the GIL for the call): Hive map tasks probing one broadcast table."""

    def __init__(self, build_keys, payload, load_factor=0.75):
        self._bt = HostTable([build_keys, payload])
        self.h = lib().orc_mj_build(C.byref(self._bt.cols[0]), None, len(build_keys),
                                    C.byref(self._bt.cols[1]), load_factor)
        if not self.h:
            raise ValueError("oracle: unsupported key type")

    def probe(self, probe_keys):
        """Inner-join matches of probe_keys: their count."""
        pt = HostTable([probe_keys])
        need = C.c_uint64(0)
        cap = max(len(probe_keys), 1)
        while True:
            rows = np.empty(cap, dtype=np.uint32)
            pay = np.empty(cap, dtype=np.int64)
            n = lib().orc_mj_probe(self.h, C.byref(pt.cols[0]), None, len(probe_keys), _ptr(rows),
                                   _ptr(pay), cap, C.byref(need))
            if n >= 0:
                return n
            cap = need.value

    def free(self):
        if self.h:
            lib().orc_mj_free(self.h)
            self.h = None


def mapjoin_stats(build_keys, load_factor=0.75):
    bt = HostTable([build_keys])
    L = lib()
    h = L.orc_mj_build(C.byref(bt.cols[0]), None, len(build_keys), None, load_factor)
    vals = [C.c_int64(), C.c_int64(), C.c_int64(), C.c_int32(), C.c_int32()]
    L.orc_mj_stats(h, *[C.byref(v) for v in vals])
    L.orc_mj_free(h)
    return dict(keys=vals[0].value, buckets=vals[1].value, values=vals[2].value,
                largest_steps=vals[3].value, expands=vals[4].value)


def join(tables, join_cols, conds):
    """N-way join -> result_offsets as an (cardinality, n_non_semi) int32 array."""
    hts = [HostTable(t) for t in tables]
    tarr = (abi.Table * len(tables))(*[h.table for h in hts])
    jc = np.ascontiguousarray(join_cols, dtype=np.uint32)
    carr = (abi.JoinCond * len(conds))(*[abi.JoinCond(*c) for c in conds])
    semi = {c[2] for c in conds if c[0] == abi.LEFT_SEMI_JOIN}
    nout = len(tables) - len(semi)
    need = C.c_uint64(0)
    cap = max(len(tables[0][0]), 1)
    while True:
        out = np.empty(cap * nout, dtype=np.int32)
        n = lib().orc_join(tarr, len(tables), _ptr(jc), carr, len(conds), _ptr(out), cap,
                           C.byref(need))
        if n == -2:
            raise ValueError("oracle: unsupported join")
        if n >= 0:
            return out[: n * nout].reshape(n, nout).copy()
        cap = need.value


def groupby(columns, keys, aggs, sel=None, exact_double_sum=True):
    """Hash group-by with Hive semantics.

    Returns (rep_rows, values, nulls): rep_rows[g] is the first row of group g (first-appearance
    order); values[a][g] the aggregate (int64 or float64 array per output type); nulls[a][g].
    """
    t = HostTable(columns)
    s = None if sel is None else np.ascontiguousarray(sel, dtype=np.uint32)
    n = len(s) if s is not None else t.table.row_num
    k = np.ascontiguousarray(keys, dtype=np.uint32)
    aarr = (abi.Agg * max(len(aggs), 1))(*[abi.Agg(a, i) for a, i in aggs])
    cap = max(n, 1)
    rep = np.empty(cap, dtype=np.uint32)
    oi = np.empty(cap * max(len(aggs), 1), dtype=np.int64)
    of = np.empty(cap * max(len(aggs), 1), dtype=np.float64)
    on = np.empty(cap * max(len(aggs), 1), dtype=np.uint8)
    g = lib().orc_groupby(t.ref(), _ptr(s), n, _ptr(k), len(keys), aarr, len(aggs),
                          1 if exact_double_sum else 0, _ptr(rep), _ptr(oi), _ptr(of), _ptr(on),
                          cap)
    if g < 0:
        raise ValueError("oracle: unsupported group-by")
    na = len(aggs)
    values, nulls = [], []
    for a, (atype, idx) in enumerate(aggs):
        itype = abi.LONG if idx == abi.COUNT_STAR else t.cols[idx].type
        otype = abi.agg_output_type(atype, itype if itype != abi.STRING else abi.LONG)
        src = of if otype in (abi.DOUBLE, abi.FLOAT) else oi
        values.append(src[a: g * na: na].copy() if na else src[:0])
        nulls.append(on[a: g * na: na].astype(bool))
    return rep[:g].copy(), values, nulls


def exact_sum(x):
    a = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_exact_sum(_ptr(a), len(a))


def sort_perm(columns, key_cols, orders):
    t = HostTable(columns)
    k = np.ascontiguousarray(key_cols, dtype=np.uint32)
    o = C.c_char_p("".join(orders).encode())
    out = np.empty(max(t.table.row_num, 1), dtype=np.uint32)
    if lib().orc_sort(t.ref(), _ptr(k), o, len(key_cols), _ptr(out)) != 0:
        raise ValueError("oracle: unsupported sort")
    return out[: t.table.row_num].copy()


def sort_runs(columns, key_cols, orders, nruns):
    """ORDER BY as Tez runs it: nruns map-side run sorts in parallel, then one merging reducer.
    Returns (permutation, seconds of the run sorts, seconds of the merge)."""
    t = HostTable(columns)
    k = np.ascontiguousarray(key_cols, dtype=np.uint32)
    o = C.c_char_p("".join(orders).encode())
    out = np.empty(max(t.table.row_num, 1), dtype=np.uint32)
    rs, ms = C.c_double(0), C.c_double(0)
    if lib().orc_sort_runs(t.ref(), _ptr(k), o, len(key_cols), nruns, _ptr(out), C.byref(rs),
                           C.byref(ms)) != 0:
        raise ValueError("oracle: unsupported sort")
    return out[: t.table.row_num].copy(), rs.value, ms.value


class FragmentSpec:
    """Host description of a CS1 fragment for orc_fragment_run.

    dims: list of dicts {columns, key_col, payload_col, fact_key_col, filter: (nodes, root)}.
    keys: list of (source, index); aggs: list of (type, expr, a, b).
    """

    def __init__(self, fact_columns, dims, keys, aggs, fact_filter=None):
        self.fact = HostTable(fact_columns)
        self.dim_tables = [HostTable(d["columns"]) for d in dims]
        self._keep = []
        self.dims = (abi.OrcDim * max(len(dims), 1))()
        for i, d in enumerate(dims):
            D = self.dims[i]
            D.table = self.dim_tables[i].table
            if d.get("filter"):
                nodes, root = d["filter"]
                arr = abi.pred_array(nodes)
                self._keep.append(arr)
                D.filter = C.cast(arr, C.POINTER(abi.Pred))
                D.filter_nnodes, D.filter_root = len(nodes), root
            D.key_col = d["key_col"]
            D.payload_col = d.get("payload_col", -1)
            D.fact_key_col = d["fact_key_col"]
        self.keys = (abi.FragmentKey * max(len(keys), 1))(*[abi.FragmentKey(*k) for k in keys])
        self.aggs = (abi.FragmentAgg * max(len(aggs), 1))(*[abi.FragmentAgg(*a) for a in aggs])
        f = abi.OrcFragment()
        f.fact = self.fact.table
        if fact_filter:
            nodes, root = fact_filter
            arr = abi.pred_array(nodes)
            self._keep.append(arr)
            f.filter = C.cast(arr, C.POINTER(abi.Pred))
            f.filter_nnodes, f.filter_root = len(nodes), root
        f.dims = C.cast(self.dims, C.POINTER(abi.OrcDim))
        f.ndims = len(dims)
        f.keys = C.cast(self.keys, C.POINTER(abi.FragmentKey))
        f.nkeys = len(keys)
        f.aggs = C.cast(self.aggs, C.POINTER(abi.FragmentAgg))
        f.naggs = len(aggs)
        self.frag = f
        self.nkeys, self.naggs = len(keys), len(aggs)

    def run(self, nthreads=1, capacity=1 << 16):
        ok = np.empty(capacity * max(self.nkeys, 1), dtype=np.int32)
        oa = np.empty(capacity * max(self.naggs, 1), dtype=np.int64)
        on = np.empty(capacity * max(self.naggs, 1), dtype=np.uint8)
        joined, secs = C.c_uint64(0), C.c_double(0)
        g = lib().orc_fragment_run(C.byref(self.frag), nthreads, _ptr(ok), _ptr(oa), _ptr(on),
                                   capacity, C.byref(joined), C.byref(secs))
        if g == -2:
            raise ValueError("oracle: unsupported fragment")
        if g < 0:
            raise ValueError("oracle: group capacity exceeded")
        keys = ok[: g * self.nkeys].reshape(g, self.nkeys).copy()
        aggs = oa[: g * self.naggs].reshape(g, self.naggs).copy()
        nulls = on[: g * self.naggs].reshape(g, self.naggs).astype(bool)
        return dict(keys=keys, aggs=aggs, nulls=nulls, joined=joined.value, seconds=secs.value)
