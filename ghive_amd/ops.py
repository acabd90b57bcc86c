"""Operator interface of the reference's gpu_impl, over libghive_mi355x.so.

Class and method names follow the reference so that tests read like its own:
FilterOperator.generatePredicate / FilterPredicate.process / FilterOperator.clean_data
(gpu_impl/src/Operator/FilterOperator.cu, FilterPredicate.cu), JoinPredicate.process
(gpu_impl/inc/Operator/BaseJoinPredicate.hpp:83), GroupByPredicate.gpu_execute
(gpu_impl/inc/Operator/GroupByPredicate.hpp:172), SinkOperator.sort (SinkOperator.cu:135-288).
Columns are torch tensors on the GPU (int32 = INT, int64 = LONG, float64 = DOUBLE,
float32 = FLOAT); torch provides device memory and the stream, the compute is the HIP library.
"""
import ctypes as C

import torch

from . import abi
from ._lib import GhiveError, lib

_TYPE_OF_DTYPE = {torch.int64: abi.LONG, torch.float64: abi.DOUBLE, torch.int32: abi.INT,
                  torch.float32: abi.FLOAT}
_DTYPE_OF_TYPE = {v: k for k, v in _TYPE_OF_DTYPE.items()}


class Context:
    """One device context (ghive_init). Work is ordered on torch's current stream."""

    def __init__(self, device=0, mem_limit=0):
        self.device = device
        h = C.c_void_p()
        st = lib().ghive_init(device, mem_limit, C.byref(h))
        if st != abi.OK:
            raise GhiveError(st, "ghive_init failed")
        self.h = h
        torch.cuda.set_device(device)
        self.use_stream(torch.cuda.current_stream(device))

    def use_stream(self, stream):
        self.check(lib().ghive_set_stream(self.h, C.c_void_p(stream.cuda_stream)))

    def check(self, status):
        if status != abi.OK:
            raise GhiveError(status, lib().ghive_last_error(self.h).decode(errors="replace"))

    def close(self):
        if self.h:
            lib().ghive_shutdown(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def device_info(self):
        d = abi.DeviceInfo()
        self.check(lib().ghive_device_query(self.h, C.byref(d)))
        return {f: (getattr(d, f).decode() if isinstance(getattr(d, f), bytes) else getattr(d, f))
                for f, _ in abi.DeviceInfo._fields_}

    def mem_stats(self):
        v = [C.c_uint64() for _ in range(4)]
        self.check(lib().ghive_mem_stats(self.h, *[C.byref(x) for x in v]))
        return dict(in_use=v[0].value, peak=v[1].value, limit=v[2].value, reserved=v[3].value)

    def set_option(self, name, value):
        self.check(lib().ghive_set_option(self.h, name.encode(), int(value)))

    def get_option(self, name):
        v = C.c_int64()
        self.check(lib().ghive_get_option(self.h, name.encode(), C.byref(v)))
        return v.value

    def stage_times(self):
        t = abi.StageTimes()
        self.check(lib().ghive_last_stage_times(self.h, C.byref(t)))
        return dict(total_ms=t.total_ms, main_kernel_ms=t.main_kernel_ms, finalize_ms=t.finalize_ms,
                    rows_in=t.rows_in, rows_joined=t.rows_joined)

    def plan_info(self):
        """What the last Plan.execute on this context did (ghive_last_plan_info)."""
        i = abi.PlanExecInfo()
        self.check(lib().ghive_last_plan_info(self.h, C.byref(i)))
        return dict(fused=bool(i.fused), joins=i.joins, narrowed_columns=i.narrowed_columns,
                    has_filter=bool(i.has_filter), fused_kernel_ms=i.fused_kernel_ms,
                    not_fused=i.not_fused.decode(errors="replace"))

    # ---- synthetic data (Philox4x32-10) -----------------------------------------------------
    def gen_uniform_i32(self, seed, col, n, lo, hi, row_begin=0, out=None):
        out = torch.empty(n, dtype=torch.int32, device=f"cuda:{self.device}") if out is None else out
        self.check(lib().ghive_gen_uniform_i32(self.h, seed, col, row_begin, n, lo, hi, _p(out)))
        return out

    def gen_uniform_i64(self, seed, col, n, lo, hi, row_begin=0):
        out = torch.empty(n, dtype=torch.int64, device=f"cuda:{self.device}")
        self.check(lib().ghive_gen_uniform_i64(self.h, seed, col, row_begin, n, lo, hi, _p(out)))
        return out

    def gen_sequence_i32(self, n, start, row_begin=0):
        out = torch.empty(n, dtype=torch.int32, device=f"cuda:{self.device}")
        self.check(lib().ghive_gen_sequence_i32(self.h, row_begin, n, start, _p(out)))
        return out

    def gen_uniform_f32(self, seed, col, n, row_begin=0):
        out = torch.empty(n, dtype=torch.float32, device=f"cuda:{self.device}")
        self.check(lib().ghive_gen_uniform_f32(self.h, seed, col, row_begin, n, _p(out)))
        return out

    def gen_rank_keys(self, seed, n, rank_begin=1):
        out = torch.empty(n, dtype=torch.int64, device=f"cuda:{self.device}")
        self.check(lib().ghive_gen_rank_keys(self.h, seed, rank_begin, n, _p(out)))
        return out

    def gen_zipf_keys(self, seed, col, n, nranks, theta=1.0, miss_frac=0.1, row_begin=0):
        out = torch.empty(n, dtype=torch.int64, device=f"cuda:{self.device}")
        self.check(lib().ghive_gen_zipf_keys(self.h, seed, col, row_begin, n, nranks, theta, miss_frac, _p(out)))
        return out

    def gen_normal_f64(self, seed, col, n, row_begin=0):
        out = torch.empty(n, dtype=torch.float64, device=f"cuda:{self.device}")
        self.check(lib().ghive_gen_normal_f64(self.h, seed, col, row_begin, n, _p(out)))
        return out


def _p(t):
    return C.c_void_p(t.data_ptr()) if t is not None and t.numel() else None


class StringColumn:
    """STRING column (gpu_impl/inc/DataFlow/Column.hpp:13-35, ghive/GTable.java:231-259): a UTF-8
    character blob (uint8 tensor) and int32 (start,end) pairs per row, (-1,-1) = NULL."""

    def __init__(self, blob, idx):
        self.blob = blob.contiguous()
        self.idx = idx.contiguous().view(-1)
        if self.blob.dtype != torch.uint8 or self.idx.dtype != torch.int32 or self.idx.numel() % 2:
            raise TypeError("STRING column: uint8 blob and int32 (start,end) pairs")
        if self.blob.numel() == 0:  # the library wants a blob address even when every string is empty
            self.blob = torch.zeros(1, dtype=torch.uint8, device=self.blob.device)
        self.is_cuda = self.blob.is_cuda
        self.device = self.blob.device

    @staticmethod
    def from_list(values, device="cuda"):
        """values: str / bytes / None (NULL)."""
        blob, idx, o = bytearray(), [], 0
        for v in values:
            if v is None:
                idx += [-1, -1]
                continue
            b = v.encode() if isinstance(v, str) else bytes(v)
            blob += b
            idx += [o, o + len(b)]
            o += len(b)
        return StringColumn(torch.tensor(list(blob), dtype=torch.uint8, device=device),
                            torch.tensor(idx, dtype=torch.int32, device=device))

    def numel(self):
        return self.idx.numel() // 2

    def contiguous(self):
        return self

    def to_list(self):
        """Host list of bytes (None for NULL)."""
        blob, idx = bytes(self.blob.cpu().numpy().tobytes()), self.idx.cpu().numpy()
        return [None if idx[2 * i] < 0 else blob[idx[2 * i]: idx[2 * i + 1]] for i in range(len(idx) // 2)]


class Table:
    """Table of device columns (gpu_impl/inc/DataFlow/Table.hpp): torch tensors or StringColumns."""

    def __init__(self, columns, key_num=0, host=False):
        """host=True: columns in host memory, for serving requests with host I/O."""
        self.columns = [c.contiguous() for c in columns]
        n = len(self.columns)
        rows = self.columns[0].numel() if n else 0
        for c in self.columns:
            if c.numel() != rows:
                raise ValueError("ragged table")
            if not isinstance(c, StringColumn) and c.dtype not in _TYPE_OF_DTYPE:
                raise TypeError(f"unsupported column dtype {c.dtype}")
            if c.is_cuda == host:
                raise ValueError("columns must be in host memory" if host else "columns must be on the GPU")
        self.host = host
        self.row_num = rows
        self.cols = (abi.Column * max(n, 1))()
        for i, c in enumerate(self.columns):
            if isinstance(c, StringColumn):
                self.cols[i].type = abi.STRING
                self.cols[i].data = c.blob.data_ptr()
                self.cols[i].aux = c.idx.data_ptr() if rows else None
                self.cols[i].char_size = c.blob.numel()
            else:
                self.cols[i].type = _TYPE_OF_DTYPE[c.dtype]
                self.cols[i].data = c.data_ptr() if c.numel() else None
            self.cols[i].row_num = rows
        self.table = abi.Table(self.cols, n, key_num, rows)

    def ref(self):
        return C.byref(self.table)


class FilterPredicate:
    """A predicate tree (FilterPredicate fields, gpu_impl/inc/Operator/FilterPredicate.hpp:61-90)."""

    def __init__(self, nodes, root):
        self.nodes = list(nodes)
        self.root = root
        self._arr = abi.pred_array(self.nodes)

    def toString(self, level=0):
        assert level == 0
        n = lib().ghive_predicate_to_string(self._arr, len(self.nodes), self.root, None, 0)
        buf = C.create_string_buffer(n)
        lib().ghive_predicate_to_string(self._arr, len(self.nodes), self.root, buf, n)
        return buf.value.decode()

    def process(self, ctx, table):
        """Ids of the rows that pass, ascending (int32 tensor holding uint32 row ids)."""
        out = torch.empty(max(table.row_num, 1), dtype=torch.int32, device=table.columns[0].device
                          if table.columns else "cuda")
        cnt = C.c_uint64()
        ctx.check(lib().ghive_filter(ctx.h, table.ref(), self._arr, len(self.nodes), self.root,
                                     _p(out), C.byref(cnt)))
        return out[: cnt.value]

    def process_async(self, ctx, table):
        """(ids, count): the passing rows' ids in the first `count` entries of a row_num-long
        tensor and their count as a 1-element device tensor, with no wait on the host."""
        dev = table.columns[0].device if table.columns else "cuda"
        out = torch.empty(max(table.row_num, 1), dtype=torch.int32, device=dev)
        cnt = torch.empty(1, dtype=torch.int64, device=dev)
        ctx.check(lib().ghive_filter_async(ctx.h, table.ref(), self._arr, len(self.nodes), self.root,
                                           _p(out), _p(cnt)))
        return out, cnt


def filtered_hash_tables(ctx, builds):
    """Hash tables of several build sides, each filtered and built in one pass
    (ghive_hash_build_filtered): builds = [(table, FilterPredicate or None, key_col, payload_col or
    None, load_factor)]. Equal to FilterPredicate.process_async then HashTable(..., count=...) per
    build; the statistics stay on the device until a table's first use."""
    n = len(builds)
    if not 1 <= n <= abi.MAX_FILTERED_BUILDS:
        raise ValueError(f"{n} builds (1..{abi.MAX_FILTERED_BUILDS})")
    jobs = (abi.FilteredBuild * n)()
    keep = []
    for i, (table, pred, key_col, pay_col, lf) in enumerate(builds):
        jobs[i].table = C.addressof(table.table)
        jobs[i].nodes = C.addressof(pred._arr) if pred is not None and pred.nodes else None
        jobs[i].nnodes = len(pred.nodes) if pred is not None else 0
        jobs[i].root = pred.root if pred is not None else 0
        jobs[i].key_col = key_col
        jobs[i].payload_col = -1 if pay_col is None else pay_col
        jobs[i].load_factor = lf
        keep.append((table, pred))
    hs = (C.c_void_p * n)()
    ctx.check(lib().ghive_hash_build_filtered(ctx.h, jobs, n, hs))
    return [HashTable(ctx, keep[i], _handle=C.c_void_p(hs[i])) for i in range(n)]


class FilterOperator:
    """FilterOperator: EXPLAIN predicate text + the operator's output columns."""

    def __init__(self, predicate_literal, output_cols):
        self.predicatesLiteral = predicate_literal
        self.output_cols = list(output_cols)

    def generatePredicate(self):
        names = (C.c_char_p * max(len(self.output_cols), 1))(*[c.encode() for c in self.output_cols])
        cap = 64
        nodes = (abi.Pred * cap)()
        nn, root = C.c_uint32(), C.c_uint32()
        err = C.create_string_buffer(512)
        st = lib().ghive_parse_predicate(self.predicatesLiteral.encode(), names, len(self.output_cols),
                                         nodes, cap, C.byref(nn), C.byref(root), err, 512)
        if st != abi.OK:
            raise GhiveError(st, err.value.decode())
        return FilterPredicate([nodes[i] for i in range(nn.value)], root.value)

    def clean_data(self, ctx, table, predicate=None):
        """Table compacted to the passing rows (the input is not modified)."""
        pred = predicate or self.generatePredicate()
        outs = [torch.empty_like(c.idx) if isinstance(c, StringColumn) else torch.empty_like(c)
                for c in table.columns]
        ocols = (abi.Column * max(len(outs), 1))()
        for i, o in enumerate(outs):
            if isinstance(table.columns[i], StringColumn):
                ocols[i].aux = o.data_ptr() if o.numel() else None
            else:
                ocols[i].data = o.data_ptr() if o.numel() else None
        cnt = C.c_uint64()
        ctx.check(lib().ghive_filter_compact(ctx.h, table.ref(), pred._arr, len(pred.nodes), pred.root,
                                             ocols, C.byref(cnt)))
        return Table([StringColumn(c.blob, o[: 2 * cnt.value]) if isinstance(c, StringColumn) else o[: cnt.value]
                      for c, o in zip(table.columns, outs)])


class HashTable:
    """Device hash table of a build side (unique or duplicate keys)."""

    def __init__(self, ctx, key, sel=None, payload=None, load_factor=0.5, count=None, _handle=None):
        """count: a 1-element device tensor holding how many entries of sel to insert (from
        FilterPredicate.process_async): the build then leaves its statistics on the device until
        the table's first use, so that filter, build and the consumer wait once."""
        self.ctx = ctx
        if _handle is not None:  # built elsewhere (filtered_hash_tables); key: what the table keeps alive
            self._keep = key
            self.h = _handle
            return
        kt = Table([key] + ([payload] if payload is not None else []))
        self._keep = (kt, key, payload, sel, count)
        n = sel.numel() if sel is not None else key.numel()
        h = C.c_void_p()
        pcol = C.byref(kt.cols[1]) if payload is not None else None
        if count is not None:
            ctx.check(lib().ghive_hash_build_async(ctx.h, C.byref(kt.cols[0]), _p(sel), _p(count), n, pcol,
                                                   load_factor, C.byref(h)))
        else:
            ctx.check(lib().ghive_hash_build(ctx.h, C.byref(kt.cols[0]), _p(sel), n, pcol, load_factor,
                                             C.byref(h)))
        self.h = h

    def info(self):
        e, s, b, u = C.c_uint64(), C.c_uint64(), C.c_uint64(), C.c_int32()
        self.ctx.check(lib().ghive_hash_table_info(self.h, C.byref(e), C.byref(s), C.byref(b), C.byref(u)))
        return dict(entries=e.value, slots=s.value, bytes=b.value, unique=bool(u.value))

    def probe(self, key, sel=None, capacity=None):
        """(probe row ids, payloads) of every match, in unspecified order."""
        kt = Table([key])
        n = sel.numel() if sel is not None else key.numel()
        cap = capacity if capacity is not None else max(n, 1)
        while True:
            rows = torch.empty(max(cap, 1), dtype=torch.int32, device=key.device)
            pay = torch.empty(max(cap, 1), dtype=torch.int64, device=key.device)
            cnt = C.c_uint64()
            st = lib().ghive_hash_probe(self.ctx.h, self.h, C.byref(kt.cols[0]), _p(sel), n, _p(rows),
                                        _p(pay), cap, C.byref(cnt))
            if st == abi.ERR_CAPACITY and capacity is None:
                cap = cnt.value
                continue
            self.ctx.check(st)
            return rows[: cnt.value], pay[: cnt.value]

    def free(self):
        if self.h:
            lib().ghive_hash_table_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class JoinPredicate:
    """BaseJoinPredicate (join_cols, join_types as (type, left, right), maintain_cols per input
    table as JoinOperator::parseExtended fills them); hash joins."""

    def __init__(self, join_cols, join_types, maintain_cols=None):
        self.join_cols = list(join_cols)
        self.join_types = [tuple(t) for t in join_types]
        self.maintain_cols = [list(m) for m in maintain_cols] if maintain_cols is not None else None
        self.result_offsets = None
        self.cardinality = 0

    def generate_result(self, ctx, tables):
        """BaseJoinPredicate::generate_result: the maintained columns of every non-semi input,
        gathered by result_offsets (NULL sentinel where an offset is -1). Call after process()."""
        if self.result_offsets is None:
            raise ValueError("generate_result before process")
        maintain = self.maintain_cols or [[] for _ in tables]
        tarr = (abi.Table * len(tables))(*[t.table for t in tables])
        conds = (abi.JoinCond * max(len(self.join_types), 1))(*[abi.JoinCond(*c) for c in self.join_types])
        rows = [(C.c_uint32 * max(len(m), 1))(*m) for m in maintain]
        mptr = (C.POINTER(C.c_uint32) * len(tables))(*[C.cast(r, C.POINTER(C.c_uint32)) for r in rows])
        mcnt = (C.c_uint32 * len(tables))(*[len(m) for m in maintain])
        semi = {t[2] for t in self.join_types if t[0] == abi.LEFT_SEMI_JOIN}
        plan = []
        for ti, t in enumerate(tables):
            if ti in semi:
                continue
            cols = maintain[ti] if maintain[ti] or plan else [0]
            plan += [t.columns[c] for c in cols]
        n = self.cardinality
        outs = [torch.empty(max(n, 1), dtype=c.dtype, device=c.device) for c in plan]
        ocols = (abi.Column * max(len(outs), 1))()
        for i, o in enumerate(outs):
            ocols[i].data = o.data_ptr()
        nout = C.c_uint32()
        ctx.check(lib().ghive_join_result(ctx.h, tarr, len(tables), conds, len(self.join_types),
                                          C.cast(mptr, C.c_void_p), C.cast(mcnt, C.c_void_p),
                                          _p(self.result_offsets), n, ocols, len(outs), C.byref(nout)))
        return [o[:n] for o in outs]

    def process(self, ctx, tables, capacity=None):
        tarr = (abi.Table * len(tables))(*[t.table for t in tables])
        jc = torch.tensor(self.join_cols, dtype=torch.int32)
        jc_arr = (C.c_uint32 * len(self.join_cols))(*self.join_cols)
        conds = (abi.JoinCond * len(self.join_types))(*[abi.JoinCond(*c) for c in self.join_types])
        nout = len(tables) - sum(1 for t in self.join_types if t[0] == abi.LEFT_SEMI_JOIN)
        cap = capacity if capacity is not None else 1
        dev = tables[0].columns[0].device
        while True:
            out = torch.empty(max(cap * nout, 1), dtype=torch.int32, device=dev)
            card = C.c_uint64()
            st = lib().ghive_join(ctx.h, tarr, len(tables), C.cast(jc_arr, C.c_void_p), conds,
                                  len(self.join_types), _p(out), cap, C.byref(card))
            if st == abi.ERR_CAPACITY and capacity is None:
                cap = card.value
                continue
            ctx.check(st)
            del jc
            self.cardinality = card.value
            self.result_offsets = out[: card.value * nout].view(card.value, nout)
            return self.result_offsets


class GroupByPredicate:
    """GroupByPredicate(keys, aggregation_descs) with Hive output types."""

    def __init__(self, keys, aggregation_descs):
        self.keys = list(keys)
        self.aggregation_descs = [tuple(a) for a in aggregation_descs]

    def output_types(self, table):
        types = [table.cols[k].type for k in self.keys]
        for t, i in self.aggregation_descs:
            it = abi.LONG if i == abi.COUNT_STAR else table.cols[i].type
            types.append(lib().ghive_agg_output_type(t, it))
        return types

    def gpu_execute(self, ctx, table, sel=None, capacity=None):
        """List of outputs: keys then aggregates, one row per group (unordered). A STRING key comes
        back as a StringColumn sharing the input's blob."""
        n = sel.numel() if sel is not None else table.row_num
        cap = capacity if capacity is not None else max(min(n, 1 << 20), 1)
        types = self.output_types(table)
        dev = table.columns[0].blob.device if table.columns and isinstance(table.columns[0], StringColumn) \
            else (table.columns[0].device if table.columns else torch.device("cuda"))
        karr = (C.c_uint32 * max(len(self.keys), 1))(*self.keys)
        aarr = (abi.Agg * max(len(self.aggregation_descs), 1))(*[abi.Agg(*a) for a in self.aggregation_descs])
        while True:
            outs = [torch.empty(2 * max(cap, 1), dtype=torch.int32, device=dev) if t == abi.STRING
                    else torch.empty(max(cap, 1), dtype=_DTYPE_OF_TYPE[t], device=dev) for t in types]
            ocols = (abi.Column * max(len(outs), 1))()
            for i, o in enumerate(outs):
                if types[i] == abi.STRING:
                    ocols[i].aux = o.data_ptr()
                else:
                    ocols[i].data = o.data_ptr()
            g = C.c_uint64()
            st = lib().ghive_groupby(ctx.h, table.ref(), _p(sel), n, C.cast(karr, C.c_void_p), len(self.keys),
                                     aarr, len(self.aggregation_descs), ocols, cap, C.byref(g))
            if st == abi.ERR_CAPACITY and capacity is None and cap < max(n, 1):
                cap = max(n, 1)
                continue
            ctx.check(st)
            res = []
            for i, o in enumerate(outs):
                if types[i] == abi.STRING:
                    res.append(StringColumn(table.columns[self.keys[i]].blob, o[: 2 * g.value]))
                else:
                    res.append(o[: g.value])
            return res


def sort(ctx, table, key_cols, orders):
    """SinkOperator ORDER BY: permutation of row ids (int32 tensor) in key order, ties by row id."""
    perm = torch.empty(max(table.row_num, 1), dtype=torch.int32, device=table.columns[0].device)
    karr = (C.c_uint32 * max(len(key_cols), 1))(*key_cols)
    ctx.check(lib().ghive_sort(ctx.h, table.ref(), C.cast(karr, C.c_void_p), "".join(orders).encode(),
                               len(key_cols), _p(perm)))
    return perm[: table.row_num]


def sort_info(ctx):
    """What the last sort on ctx did (ghive_last_sort_info): key words, passes, bytes moved."""
    t = abi.SortInfo()
    ctx.check(lib().ghive_last_sort_info(ctx.h, C.byref(t)))
    return {f: getattr(t, f) for f, _ in abi.SortInfo._fields_ if f != "reserved"}


def _column_of(t):
    c = abi.Column()
    c.type = _TYPE_OF_DTYPE[t.dtype]
    c.data = t.data_ptr() if t.numel() else None
    c.row_num = t.numel()
    return c


def select_arith(ctx, a, b, op):
    """CalculationSelectPredicate::process (gpu_impl/src/Operator/SelectPredicate.cu:84-409):
    a op b elementwise over device tensors (a 1-element tensor broadcasts). The result dtype is
    float64 for '/' or a floating operand, int32 for int32 op int32, else int64."""
    if not (a.is_cuda and b.is_cuda):
        raise ValueError("operands must be on the GPU")
    a, b = a.contiguous(), b.contiguous()
    n = max(a.numel(), b.numel()) if min(a.numel(), b.numel()) == 1 else a.numel()
    fp = a.dtype in (torch.float32, torch.float64) or b.dtype in (torch.float32, torch.float64)
    dt = torch.float64 if op == "/" or fp else (torch.int32 if a.dtype == b.dtype == torch.int32 else torch.int64)
    out = torch.empty(n, dtype=dt, device=a.device)
    ca, cb, co = _column_of(a), _column_of(b), _column_of(out)
    ctx.check(lib().ghive_select_arith(ctx.h, C.byref(ca), C.byref(cb), op.encode(), n, C.byref(co)))
    return out


def _any_column_of(c):
    if isinstance(c, StringColumn):
        col = abi.Column()
        col.type = abi.STRING
        col.data = c.blob.data_ptr()
        col.aux = c.idx.data_ptr() if c.numel() else None
        col.row_num = c.numel()
        col.char_size = c.blob.numel()
        return col
    return _column_of(c)


def select_cast(ctx, x, to_type):
    """CAST of a device column (tensor or StringColumn) to INT / LONG / DOUBLE / FLOAT, as Hive's
    UDFToInteger / UDFToLong / UDFToDouble (string_convert_int_functor, SelectPredicate.hpp:22-66)."""
    n = x.numel()
    dev = x.device
    out = torch.empty(max(n, 1), dtype=_DTYPE_OF_TYPE[to_type], device=dev)
    ci, co = _any_column_of(x), _column_of(out)
    ctx.check(lib().ghive_select_cast(ctx.h, C.byref(ci), to_type, n, C.byref(co)))
    return out[:n]


def select_substr(ctx, s, pos, length):
    """SUBSTR(s, pos, length) of a StringColumn as Hive's UDFSubstr; the result shares s's blob."""
    n = s.numel()
    pairs = torch.empty(max(2 * n, 2), dtype=torch.int32, device=s.device)
    ci, co = _any_column_of(s), abi.Column()
    co.aux = pairs.data_ptr()
    ctx.check(lib().ghive_select_substr(ctx.h, C.byref(ci), pos, length, n, C.byref(co)))
    return StringColumn(s.blob, pairs[: 2 * n])


class Fragment:
    """Fused CS1 fragment: fact table, star joins (fact column, HashTable), keys, aggregates."""

    def __init__(self, fact, joins, keys, aggs, fact_filter=None):
        self.fact = fact
        self.joins_py = list(joins)
        self.jarr = (abi.FragmentJoin * max(len(joins), 1))(
            *[abi.FragmentJoin(c, ht.h.value) for c, ht in joins])
        self.karr = (abi.FragmentKey * max(len(keys), 1))(*[abi.FragmentKey(*k) for k in keys])
        self.aarr = (abi.FragmentAgg * max(len(aggs), 1))(*[abi.FragmentAgg(*a) for a in aggs])
        f = abi.Fragment()
        f.fact = fact.table
        if fact_filter is not None:
            f.filter = C.cast(fact_filter._arr, C.POINTER(abi.Pred))
            f.filter_nnodes, f.filter_root = len(fact_filter.nodes), fact_filter.root
            self._filter = fact_filter
        f.joins = C.cast(self.jarr, C.POINTER(abi.FragmentJoin))
        f.njoins = len(joins)
        f.keys = C.cast(self.karr, C.POINTER(abi.FragmentKey))
        f.nkeys = len(keys)
        f.aggs = C.cast(self.aarr, C.POINTER(abi.FragmentAgg))
        f.naggs = len(aggs)
        self.f = f
        self.nkeys, self.naggs = len(keys), len(aggs)

    def rebind(self, tables):
        """The same fragment over new dimension tables (one HashTable per join, in join order)."""
        if len(tables) != len(self.joins_py):
            raise ValueError("one table per join")
        self.joins_py = [(c, ht) for (c, _), ht in zip(self.joins_py, tables)]
        for i, ht in enumerate(tables):
            self.jarr[i].table = ht.h.value

    def execute(self, ctx, capacity=4096, late=False, outs=None, noalloc=False):  # noqa: D102 (see class doc)
        return self._execute(ctx, capacity, late, outs, noalloc)

    def _execute(self, ctx, capacity, late, outs, noalloc=False):
        dev = self.fact.columns[0].device
        if outs is None:
            outs = ([torch.empty(capacity, dtype=torch.int32, device=dev) for _ in range(self.nkeys)] +
                    [torch.empty(capacity, dtype=torch.int64, device=dev) for _ in range(self.naggs)])
        ocols = (abi.Column * len(outs))()
        for i, o in enumerate(outs):
            ocols[i].data = o.data_ptr()
        g = C.c_uint64()
        fn = (lib().ghive_fragment_execute_noalloc if noalloc else
              lib().ghive_fragment_execute_late if late else lib().ghive_fragment_execute)
        ctx.check(fn(ctx.h, C.byref(self.f), ocols, capacity, C.byref(g)))
        # aggregates over DOUBLE inputs come back as DOUBLE in the 8-byte buffers
        return [(o.view(torch.float64) if ocols[i].type == abi.DOUBLE and o.dtype == torch.int64 else o)[: g.value]
                for i, o in enumerate(outs)]


class Batch:
    """Batched form (ghive_batch_*): many fragments of one query shape in one persistent kernel
    replayed from a HIP graph. items: list of (key, a, b, filter_lo, filter_hi) where key/a/b
    are int32 device tensors of one fragment (b may be None)."""

    def __init__(self, ctx, dim, items, filter_shift, filter_bits, group_shift, group_bits,
                 tile_rows=0, blocks_per_cu=0, probe_mode=0, use_graph=True):
        self.ctx = ctx
        self._keep = [dim, items]
        self.spec = abi.BatchSpec(dim.h.value, filter_shift, filter_bits, group_shift, group_bits)
        arr = (abi.BatchItem * max(len(items), 1))()
        for i, (k, a, b, lo, hi) in enumerate(items):
            arr[i] = abi.BatchItem(k.data_ptr() if k.numel() else None, a.data_ptr() if a.numel() else None,
                                   b.data_ptr() if b is not None and b.numel() else None, k.numel(), lo, hi)
        self.n = len(items)
        tn = abi.BatchTuning(tile_rows, blocks_per_cu, probe_mode, 0)
        h = C.c_void_p()
        ctx.check(lib().ghive_batch_create(ctx.h, C.byref(self.spec), arr, len(items), C.byref(tn),
                                           1 if use_graph else 0, C.byref(h)))
        self.h = h
        self.groups = 1 << group_bits

    def run(self):
        """Runs the batch; returns the device time in ms (HIP events around the launch)."""
        self.ctx.check(lib().ghive_batch_run(self.ctx.h, self.h, C.byref(self.spec)))
        return self.info()["last_ms"]

    def results(self):
        import numpy as np
        sums = np.empty(self.n * self.groups, dtype=np.int64)
        counts = np.empty(self.n * self.groups, dtype=np.int64)
        self.ctx.check(lib().ghive_batch_results(self.ctx.h, self.h, sums.ctypes.data_as(C.c_void_p),
                                                 counts.ctypes.data_as(C.c_void_p)))
        return sums.reshape(self.n, self.groups), counts.reshape(self.n, self.groups)

    def info(self):
        g, rows, tiles, mode, grid, ms = (C.c_uint32(), C.c_uint64(), C.c_uint64(), C.c_int32(),
                                          C.c_uint32(), C.c_float())
        self.ctx.check(lib().ghive_batch_info(self.h, C.byref(g), C.byref(rows), C.byref(tiles),
                                              C.byref(mode), C.byref(grid), C.byref(ms)))
        return dict(groups=g.value, rows=rows.value, tiles=tiles.value, probe_mode=mode.value,
                    grid=grid.value, last_ms=ms.value)

    def free(self):
        if self.h:
            lib().ghive_batch_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def bench_stream(ctx, nbytes, mode=0, iters=10):
    """Achievable HBM bandwidth (GB/s): mode 0 read, mode 1 copy (counts read + write), mode 2 read
    as five columns streamed side by side (nbytes a multiple of 80)."""
    src = torch.empty(nbytes // 4, dtype=torch.int32, device=f"cuda:{ctx.device}").fill_(1)
    dst = torch.empty_like(src) if mode == 1 else None
    ms = C.c_float()
    ctx.check(lib().ghive_bench_stream(ctx.h, _p(dst), _p(src), nbytes, mode, iters, C.byref(ms)))
    moved = nbytes * (2 if mode == 1 else 1)
    return moved / (ms.value * 1e-3) / 1e9, ms.value


class Ticket:
    """A submitted serving request: keeps its descriptors and output buffers alive until waited."""

    def __init__(self, rid, kind, outs, keep):
        self.id, self.kind, self.outs, self._keep = rid, kind, outs, keep
        self.response = None


class Server:
    """Serving form (ghive_server_*): a request loop with `slots` concurrent requests on one
    device, admission by HBM footprint (inputs + intermediates) against `hbm_budget` bytes, and
    resident dimension hash tables (HashTable objects built by the caller and shared by every
    request). Inputs of device requests must be complete when submitted (the slots run on their
    own streams); host_io requests take host tables and return host tensors."""

    def __init__(self, device=0, slots=4, hbm_budget=64 << 30, slot_mem_limit=16 << 30):
        cfg = abi.ServerConfig(slots, 0, hbm_budget, slot_mem_limit)
        h = C.c_void_p()
        st = lib().ghive_server_create(device, C.byref(cfg), C.byref(h))
        if st != abi.OK:
            raise GhiveError(st, f"ghive_server_create failed with status {st}")
        self.h, self.device, self.slots = h, device, slots

    def set_option(self, name, value):
        for i in range(self.slots):
            c = lib().ghive_server_slot_context(self.h, i)
            st = lib().ghive_set_option(c, name.encode(), value)
            if st != abi.OK:
                raise GhiveError(st, lib().ghive_last_error(c).decode())

    def _submit(self, req, kind, outs, keep):
        if not req.host_io and torch.cuda.is_available():
            torch.cuda.current_stream(self.device).synchronize()
        rid = C.c_uint64()
        st = lib().ghive_server_submit(self.h, C.byref(req), C.byref(rid))
        if st != abi.OK:
            raise GhiveError(st, "request rejected: its HBM footprint exceeds the admission budget"
                             if st == abi.ERR_ADMISSION else f"submit failed with status {st}")
        return Ticket(rid.value, kind, outs, keep + [req])

    def _outs(self, dtypes, rows, host):
        if host:
            return [torch.empty(max(rows, 1), dtype=d, pin_memory=True) for d in dtypes]
        return [torch.empty(max(rows, 1), dtype=d, device=f"cuda:{self.device}") for d in dtypes]

    def submit_fragment(self, frag, capacity=4096, footprint=0):
        host = getattr(frag.fact, "host", False)
        outs = self._outs([torch.int32] * frag.nkeys + [torch.int64] * frag.naggs, capacity, host)
        ocols = (abi.Column * max(len(outs), 1))()
        for i, o in enumerate(outs):
            ocols[i].type = _TYPE_OF_DTYPE[o.dtype]
            ocols[i].data = o.data_ptr()
        req = abi.Request(kind=abi.REQ_FRAGMENT, host_io=int(host), fragment=C.pointer(frag.f),
                          out_cols=ocols, capacity=capacity, footprint_bytes=footprint)
        return self._submit(req, abi.REQ_FRAGMENT, outs, [frag, ocols])

    def submit_groupby(self, table, keys, aggregation_descs, capacity, footprint=0):
        gb = GroupByPredicate(keys, aggregation_descs)
        outs = self._outs([_DTYPE_OF_TYPE[t] for t in gb.output_types(table)], capacity, table.host)
        ocols = (abi.Column * max(len(outs), 1))()
        for i, o in enumerate(outs):
            ocols[i].type = _TYPE_OF_DTYPE[o.dtype]
            ocols[i].data = o.data_ptr()
        karr = (C.c_uint32 * max(len(gb.keys), 1))(*gb.keys)
        aarr = (abi.Agg * max(len(gb.aggregation_descs), 1))(*[abi.Agg(*a) for a in gb.aggregation_descs])
        req = abi.Request(kind=abi.REQ_GROUPBY, host_io=int(table.host), table=C.pointer(table.table),
                          keys=karr, nkeys=len(gb.keys), naggs=len(gb.aggregation_descs), aggs=aarr,
                          out_cols=ocols, capacity=capacity, footprint_bytes=footprint)
        return self._submit(req, abi.REQ_GROUPBY, outs, [table, ocols, karr, aarr])

    def submit_sort(self, table, key_cols, orders, footprint=0):
        (perm,) = self._outs([torch.int32], table.row_num, table.host)
        karr = (C.c_uint32 * max(len(key_cols), 1))(*key_cols)
        ords = "".join(orders).encode()
        req = abi.Request(kind=abi.REQ_SORT, host_io=int(table.host), table=C.pointer(table.table),
                          keys=karr, orders=ords, nkeys=len(key_cols), out_perm=perm.data_ptr(),
                          footprint_bytes=footprint)
        return self._submit(req, abi.REQ_SORT, [perm], [table, karr, ords])

    def wait(self, ticket):
        """Blocks until the request has finished; returns its outputs (group-by / fragment:
        key and aggregate columns cut to the group count; sort: the permutation)."""
        r = abi.Response()
        st = lib().ghive_server_wait(self.h, ticket.id, C.byref(r))
        if st != abi.OK:
            raise GhiveError(st, f"wait failed for request {ticket.id}")
        ticket.response = r
        if r.status != abi.OK:
            raise GhiveError(r.status, r.error.decode(errors="replace"))
        return [o[: r.groups] for o in ticket.outs]

    def stats(self):
        s = abi.ServerStats()
        lib().ghive_server_get_stats(self.h, C.byref(s))
        return {f: getattr(s, f) for f, _ in abi.ServerStats._fields_}

    def close(self):
        if self.h:
            lib().ghive_server_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gtables_struct(tables, key_cnt=None):
    """ghive_gtables for a list of (vertex name or None, [columns]) -- a column is a numpy array
    (int64 LONG, float64 DOUBLE, int32 INT) or a (bytes blob, int32 (start,end) array) STRING column --
    with the buffers of each type concatenated across tables in order, as the Java side hands them
    over. Returns (struct, objects that must stay alive while it is used)."""
    import numpy as np
    keep = []
    typecode = {np.dtype(np.int64): 0, np.dtype(np.float64): 1, np.dtype(np.int32): 2}
    bufs = {0: [], 1: [], 2: [], 3: [], "idx": []}
    types, ncols, rows, names = [], [], [], []
    for name, cols in tables:
        tt = []
        n = None
        for c in cols:
            if isinstance(c, tuple):
                blob, idx = c
                idx = np.ascontiguousarray(idx, dtype=np.int32)
                b = C.create_string_buffer(bytes(blob), max(len(blob), 1))
                keep += [b, idx]
                bufs[3].append(C.cast(b, C.c_void_p))
                bufs["idx"].append(idx.ctypes.data)
                tt.append(3)
                n = len(idx) // 2
            else:
                a = np.ascontiguousarray(c)
                keep.append(a)
                code = typecode[a.dtype]
                bufs[code].append(a.ctypes.data)
                tt.append(code)
                n = len(a)
        arr = (C.c_int32 * max(len(tt), 1))(*tt)
        keep.append(arr)
        types.append(C.cast(arr, C.POINTER(C.c_int32)))
        ncols.append(len(tt))
        rows.append(n or 0)
        names.append(None if name is None else name.encode())
    g = abi.GTables()
    g.ntables = len(tables)
    g.types = (C.POINTER(C.c_int32) * max(len(types), 1))(*types)
    g.ncols = (C.c_uint32 * max(len(ncols), 1))(*ncols)
    g.row_cnt = (C.c_int32 * max(len(rows), 1))(*rows)
    if key_cnt is not None:
        g.key_cnt = (C.c_int32 * max(len(key_cnt), 1))(*key_cnt)
    g.vertex_name = (C.c_char_p * max(len(names), 1))(*names)
    for code, fld, cnt in ((0, "long_cols", "nlong"), (1, "double_cols", "ndouble"), (2, "int_cols", "nint"),
                           (3, "string_cols", "nstring")):
        setattr(g, fld, (C.c_void_p * max(len(bufs[code]), 1))(*bufs[code]))
        setattr(g, cnt, len(bufs[code]))
    g.string_idx_cols = (C.c_void_p * max(len(bufs["idx"]), 1))(*bufs["idx"])
    return g, keep


class GTablesStaged:
    """GTables (ghive/GTables.java) staged through ghive_gtables_stage. tables: a list of
    (vertex name or None, [columns]) where a column is a numpy array (int64 LONG, float64 DOUBLE,
    int32 INT) or a (bytes blob, int32 (start,end) array) STRING column; the buffers of each type
    are concatenated across tables in order, as the Java side hands them over."""

    def __init__(self, ctx, tables, key_cnt=None):
        self.ctx, self.h = ctx, C.c_void_p()
        g, self._keep = gtables_struct(tables, key_cnt)
        ctx.check(lib().ghive_gtables_stage(ctx.h, C.byref(g), C.byref(self.h)))

    def inputs(self):
        p, n = C.POINTER(abi.PlanInput)(), C.c_uint32()
        lib().ghive_staged_inputs(self.h, C.byref(p), C.byref(n))
        return p, n.value

    def close(self):
        if self.h:
            lib().ghive_staged_free(self.ctx.h, self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpuresult_sequence(types):
    """generate_tbl_result's numbering of result columns: (sequence, counts per LONG/DOUBLE/INT/STRING)."""
    t = (C.c_int32 * max(len(types), 1))(*types)
    out = (C.c_int32 * max(len(types), 1))()
    counts = (C.c_uint32 * 4)()
    st = lib().ghive_gpuresult_sequence(t, len(types), out, counts)
    if st != abi.OK:
        raise GhiveError(st, "gpuresult_sequence: unknown column type")
    return list(out)[: len(types)], list(counts)


class Plan:
    """A query plan (ghive_plan_*): Hive's EXPLAIN and EXPLAIN EXTENDED texts (and the Java side's
    maintain-values JSON) parsed into every vertex's operator tree with resolved descriptors.
    describe() returns {"root", "vertices": {vertex: sink id}, "operators": {id: {...}}}.
    plan_text None: the EXPLAIN EXTENDED text alone gives the tree."""

    def __init__(self, plan_text, extended_text=None, maintain_json=None):
        import json
        err = C.create_string_buffer(512)
        h = C.c_void_p()
        st = lib().ghive_plan_parse(plan_text.encode() if plan_text else None,
                                    extended_text.encode() if extended_text else None,
                                    maintain_json.encode() if maintain_json else None, err, len(err), C.byref(h))
        if st != abi.OK:
            raise GhiveError(st, err.value.decode(errors="replace"))
        self.h = h
        n = C.c_uint64()
        lib().ghive_plan_describe(self.h, None, 0, C.byref(n))
        buf = C.create_string_buffer(n.value + 1)
        st = lib().ghive_plan_describe(self.h, buf, n.value + 1, C.byref(n))
        if st != abi.OK:
            raise GhiveError(st, "plan describe failed")
        self.desc = json.loads(buf.value.decode())

    def describe(self):
        return self.desc

    def op(self, op_id):
        return self.desc["operators"][op_id]

    def chain(self, vertex):
        """Operator ids of a vertex from its sink down to its table scan / reduce input."""
        out, node = [], self.desc["vertices"][vertex]
        while node is not None:
            out.append(node)
            kids = [c for c in self.op(node)["children"]
                    if self.op(c)["kind"] not in ("ReduceSink", "FileSink", "Union")]
            node = kids[0] if kids else None
        return out

    def execute(self, ctx, vertex, inputs):
        """Operator::execute over one vertex (ghive_plan_execute). inputs: {name: (Table, column
        names or None)} -- a base table by its TableScan name ("db@table") or by the vertex run, or
        another vertex's result by that vertex's name; None = positional columns. Returns (column
        names, [device tensors / StringColumns]); the columns are copies owned by torch."""
        if isinstance(inputs, GTablesStaged):  # tables staged from GTables buffers
            arr, count = inputs.inputs()
            keep = [inputs]
        else:
            keep, arr, count = [], (abi.PlanInput * max(len(inputs), 1))(), len(inputs)
            for i, (name, (table, names)) in enumerate(inputs.items()):
                keep.append(table)
                arr[i].name = name.encode()
                arr[i].table = C.pointer(table.table)
                if names is not None:  # else positional, as GTables delivers the columns
                    cn = (C.c_char_p * max(len(names), 1))(*[n.encode() for n in names])
                    keep.append(cn)
                    arr[i].column_names = C.cast(cn, C.POINTER(C.c_char_p))
        err = C.create_string_buffer(512)
        r = C.c_void_p()
        st = lib().ghive_plan_execute(ctx.h, self.h, vertex.encode(), arr, count, err, len(err), C.byref(r))
        if st != abi.OK:
            raise GhiveError(st, err.value.decode(errors="replace"))
        try:
            rows, ncols = C.c_uint64(), C.c_uint32()
            ctx.check(lib().ghive_result_info(r, C.byref(rows), C.byref(ncols)))
            names, cols = [], []
            for i in range(ncols.value):
                c, nm = abi.Column(), C.c_char_p()
                ctx.check(lib().ghive_result_column(r, i, C.byref(c), C.byref(nm)))
                dev = f"cuda:{ctx.device}"
                if c.type == abi.STRING:  # copy the (start,end) pairs and the blob they point into
                    idx = torch.empty(2 * rows.value, dtype=torch.int32, device=dev)
                    blob = torch.empty(max(c.char_size, 1), dtype=torch.uint8, device=dev)
                    if rows.value:
                        ctx.check(lib().ghive_memcpy(ctx.h, _p(idx), c.aux, 8 * rows.value, 2))
                    if c.char_size:
                        ctx.check(lib().ghive_memcpy(ctx.h, _p(blob), c.data, c.char_size, 2))
                    t = StringColumn(blob, idx)
                else:
                    t = torch.empty(rows.value, dtype=_DTYPE_OF_TYPE[c.type], device=dev)
                    if rows.value:
                        ctx.check(lib().ghive_memcpy(ctx.h, _p(t), c.data, rows.value * t.element_size(), 2))
                names.append(nm.value.decode())
                cols.append(t)
            return names, cols
        finally:
            lib().ghive_result_free(ctx.h, r)

    def close(self):
        if self.h:
            lib().ghive_plan_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gpu_process(ctx, plan, tables, this_vertex, key_cnt=None):
    """GPUProcess in one call (ghive_gpu_process): GTables host buffers in (as gtables_struct takes
    them), the vertex run, the GPUResult out as (column arrays in output order, sequence, counts), or
    None for a missing input (a table named None). STRING columns come back as (blob, pairs)."""
    import numpy as np
    g, keep = gtables_struct(tables, key_cnt)
    err = C.create_string_buffer(512)
    h = C.c_void_p()
    st = lib().ghive_gpu_process(ctx.h, plan.h, C.byref(g), this_vertex.encode(), err, len(err), C.byref(h))
    if st != abi.OK:
        raise GhiveError(st, err.value.decode(errors="replace") or lib().ghive_last_error(ctx.h).decode())
    if not h.value:
        return None
    try:
        rows, ncols = C.c_uint64(), C.c_uint32()
        counts = (C.c_uint32 * 4)()
        seq = C.POINTER(C.c_int32)()
        ctx.check(lib().ghive_gpu_result_info(h, C.byref(rows), C.byref(ncols), counts, C.byref(seq)))
        n = rows.value

        def arr(ptr, count, dtype):  # an empty result column has no buffer
            return np.ctypeslib.as_array(ptr, (count,)).copy() if count else np.zeros(0, dtype=dtype)

        groups = []
        for i in range(counts[0]):
            groups.append(arr(lib().ghive_gpu_result_long(h, i), n, np.int64))
        for i in range(counts[1]):
            groups.append(arr(lib().ghive_gpu_result_double(h, i), n, np.float64))
        for i in range(counts[2]):
            groups.append(arr(lib().ghive_gpu_result_int(h, i), n, np.int32))
        for i in range(counts[3]):
            blob, nb, pairs = C.c_char_p(), C.c_uint64(), C.POINTER(C.c_int32)()
            ctx.check(lib().ghive_gpu_result_string(h, i, C.byref(blob), C.byref(nb), C.byref(pairs)))
            b = C.string_at(blob, nb.value) if nb.value else b""
            groups.append((b, arr(pairs, 2 * n, np.int32)))
        sequence = [seq[c] for c in range(ncols.value)]
        cols = [groups[q] for q in sequence]
        return cols, sequence, list(counts)
    finally:
        lib().ghive_gpu_result_free(h)
        del keep


def process_times(ctx):
    """Wall-clock split of the last gpu_process (ghive_last_process_times)."""
    t = abi.ProcessTimes()
    ctx.check(lib().ghive_last_process_times(ctx.h, C.byref(t)))
    return {f: getattr(t, f) for f, _ in abi.ProcessTimes._fields_}
