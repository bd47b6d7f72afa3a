"""Python harness over the C ABI of libvj.so (include/vj.h), via ctypes.

Used by the tests and bench.py. The join and hash work runs in the HIP library;
this module only marshals descriptors. Device buffers are torch tensors on a
ROCm device (torch is used for memory and streams, nothing else).

Loading fails loudly when libvj.so is missing: there is no CPU fallback.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libvj.so")
_LIB = None

VJ_OK = 0
STATUS = {0: "VJ_OK", 1: "VJ_ERR_INVALID_ARGUMENT", 2: "VJ_ERR_UNKNOWN_HASH", 3: "VJ_ERR_UNSUPPORTED",
          4: "VJ_ERR_BUDGET", 5: "VJ_ERR_OUT_OF_MEMORY", 6: "VJ_ERR_CAPACITY", 7: "VJ_ERR_HIP",
          8: "VJ_ERR_NO_DEVICE", 9: "VJ_ERR_NOT_FOUND", 10: "VJ_ERR_INTERNAL"}
STATUS_CODE = {v: k for k, v in STATUS.items()}

HASH_KINDS = ["fnv", "custom_add", "custom_add_shift", "custom_add_hw", "custom_xor", "custom_xor_shift",
              "custom_xor_hw", "custom_mult", "custom_mult_shift", "custom_mult_hw", "custom_mult_xor_shift",
              "custom_n_mult_xor_shift", "custom_mult_xor_hw"]
WORDS_REFERENCE, WORDS_NATIVE = 0, 1


class VJError(RuntimeError):
    def __init__(self, code, where, detail=""):
        self.code = code
        self.name = STATUS.get(code, str(code))
        super().__init__("%s failed: %s %s" % (where, self.name, detail))


class Table(ctypes.Structure):
    _fields_ = [("n_rows", ctypes.c_uint64), ("n_cols", ctypes.c_int32), ("key_width", ctypes.c_int32),
                ("key_signed", ctypes.c_int32), ("reserved", ctypes.c_int32), ("cols", ctypes.c_void_p * 4),
                ("rowids", ctypes.c_void_p)]


class HashOpts(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("hash_bits", ctypes.c_int32), ("chunk_bits", ctypes.c_int32),
                ("word_mode", ctypes.c_int32)]


class JoinOpts(ctypes.Structure):
    _fields_ = [("hash", HashOpts), ("passes", ctypes.c_int32), ("bits1", ctypes.c_int32),
                ("bits2", ctypes.c_int32), ("one_based", ctypes.c_int32), ("out_capacity_hint", ctypes.c_uint64),
                ("tile", ctypes.c_int32), ("table", ctypes.c_int32), ("layout", ctypes.c_int32),
                ("reserved", ctypes.c_int32)]


# Kernel shapes (vj_join_opts.tile / .table; 0 = planner / autotuned)
TILE_STANDARD, TILE_SMALL, TILE_LARGE = 1, 2, 3
TABLE_STANDARD, TABLE_WIDE = 1, 2
# Partition layouts (vj_join_opts.layout; 0 = planner / autotuned)
LAYOUT_EXACT, LAYOUT_REGIONS = 1, 2


class Rows(ctypes.Structure):
    """vj_rows: a table in the reference layout (row-major uint64 key columns + uint64 primary keys)."""
    _fields_ = [("n_rows", ctypes.c_uint64), ("n_cols", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("column_values", ctypes.c_void_p), ("primary_keys", ctypes.c_void_p)]


class Pairs(ctypes.Structure):
    _fields_ = [("count", ctypes.c_uint64), ("build_rowids", ctypes.c_void_p), ("probe_rowids", ctypes.c_void_p)]


class Request(ctypes.Structure):
    _fields_ = [("build", ctypes.c_void_p), ("table", ctypes.c_uint64), ("probe", ctypes.c_void_p),
                ("out_build", ctypes.c_void_p), ("out_probe", ctypes.c_void_p), ("capacity", ctypes.c_uint64),
                ("keep_build", ctypes.c_int32), ("reserved", ctypes.c_int32)]


class Response(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("batch_size", ctypes.c_uint32), ("count", ctypes.c_uint64),
                ("table", ctypes.c_uint64), ("t_submit_ns", ctypes.c_uint64), ("t_launch_ns", ctypes.c_uint64),
                ("t_done_ns", ctypes.c_uint64)]


class ServerStats(ctypes.Structure):
    _fields_ = [(n, ctypes.c_uint64) for n in ("submitted", "completed", "failed", "batches", "overlapped_batches",
                                                "max_batch", "resident_tables", "resident_bytes", "peak_bytes",
                                                "budget_waits", "budget_deferred")]


class JoinSummary(ctypes.Structure):
    _fields_ = [("build_rows", ctypes.c_uint64), ("probe_rows", ctypes.c_uint64), ("matches", ctypes.c_uint64),
                ("passes", ctypes.c_int32), ("bits1", ctypes.c_int32), ("bits2", ctypes.c_int32),
                ("retries", ctypes.c_int32), ("partitions", ctypes.c_uint64), ("tasks", ctypes.c_uint64),
                ("ms_total", ctypes.c_double), ("ms_histogram", ctypes.c_double), ("ms_pass1", ctypes.c_double),
                ("ms_pass2", ctypes.c_double), ("ms_build_probe", ctypes.c_double),
                ("scratch_bytes", ctypes.c_uint64), ("graph_replay", ctypes.c_int32),
                ("layout_fallback", ctypes.c_int32),
                ("bytes_histogram", ctypes.c_uint64), ("bytes_pass1", ctypes.c_uint64),
                ("bytes_pass2", ctypes.c_uint64), ("bytes_build_probe", ctypes.c_uint64),
                ("bytes_total", ctypes.c_uint64), ("gbps_histogram", ctypes.c_double),
                ("gbps_pass1", ctypes.c_double), ("gbps_pass2", ctypes.c_double),
                ("gbps_build_probe", ctypes.c_double), ("gbps_total", ctypes.c_double),
                ("probe_tuples_per_s", ctypes.c_double), ("tile_rows", ctypes.c_int32),
                ("probe_threads", ctypes.c_int32), ("table_rows", ctypes.c_int32), ("layout", ctypes.c_int32)]


def lib():
    """Loads libvj.so (raises if it is not built)."""
    global _LIB
    if _LIB is None:
        # torch first: its bundled libamdhip64.so.7 then also serves libvj.so (same SONAME),
        # so the process has exactly one HIP runtime.
        import torch  # noqa: F401
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libvj.so not built: run __graft_entry__.build() (no CPU fallback exists)")
        L = ctypes.CDLL(LIB_PATH)
        vp, i32, u64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_uint64
        P = ctypes.POINTER
        sigs = {
            "vj_status_str": ([i32], ctypes.c_char_p),
            "vj_hash_opts_default": ([P(HashOpts)], None),
            "vj_hash_kind_from_name": ([ctypes.c_char_p, P(i32)], i32),
            "vj_hash_kind_name": ([i32], ctypes.c_char_p),
            "vj_hash": ([P(Table), P(HashOpts), vp, vp], i32),
        }
        optional = {
            "vj_ctx_create": ([ctypes.c_int, P(vp)], i32),
            "vj_ctx_destroy": ([vp], i32),
            "vj_ctx_set_budget": ([vp, u64], i32),
            "vj_ctx_memory": ([vp, P(u64)], i32),
            "vj_ctx_set_profiling": ([vp, ctypes.c_int], i32),
            "vj_last_error": ([vp], ctypes.c_char_p),
            "vj_join_opts_default": ([P(JoinOpts)], None),
            "vj_join": ([vp, P(Table), P(Table), P(JoinOpts), P(Pairs), vp], i32),
            "vj_join_into": ([vp, P(Table), P(Table), P(JoinOpts), vp, vp, u64, P(u64), vp], i32),
            "vj_join_into_noalloc": ([vp, P(Table), P(Table), P(JoinOpts), vp, vp, u64, P(u64), vp], i32),
            "vj_pairs_free": ([P(Pairs)], i32),
            "vj_join_rows": ([vp, P(Rows), P(Rows), P(JoinOpts), P(Rows), vp], i32),
            "vj_rows_free": ([P(Rows)], i32),
            "vj_join_summary_get": ([vp, P(JoinSummary)], i32),
            "vj_partition": ([vp, P(Table), P(HashOpts), i32, i32, P(Table), vp, vp, vp], i32),
            "vj_join_batch": ([vp, i32, P(Table), P(Table), P(JoinOpts), P(vp), P(vp), P(u64), P(u64), i32, vp],
                              i32),
            "vj_tune": ([vp, P(Table), P(Table), P(JoinOpts), i32, P(JoinOpts), P(ctypes.c_double), vp], i32),
            "vj_tune_clear": ([vp], i32),
            "vj_build": ([vp, P(Table), P(JoinOpts), P(vp), vp], i32),
            "vj_probe_into": ([vp, vp, P(Table), P(JoinOpts), vp, vp, u64, P(u64), vp], i32),
            "vj_built_bytes": ([vp, P(u64)], i32),
            "vj_built_free": ([vp], i32),
            "vj_server_create": ([ctypes.c_int, u64, P(JoinOpts), i32, P(vp)], i32),
            "vj_server_submit": ([vp, P(Request), P(u64)], i32),
            "vj_server_wait": ([vp, u64, P(Response)], i32),
            "vj_server_poll": ([vp, u64, P(Response), P(i32)], i32),
            "vj_server_drop_table": ([vp, u64], i32),
            "vj_server_stats_get": ([vp, P(ServerStats)], i32),
            "vj_server_destroy": ([vp], i32),
        }
        for name, (args, res) in list(sigs.items()) + list(optional.items()):
            if not hasattr(L, name):
                if name in sigs:
                    raise RuntimeError("libvj.so lacks %s" % name)
                continue
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _LIB = L
    return _LIB


def check(rc, where, ctx=None):
    if rc != VJ_OK:
        detail = ""
        if ctx is not None and hasattr(lib(), "vj_last_error"):
            msg = lib().vj_last_error(ctx)
            detail = msg.decode() if msg else ""
        raise VJError(rc, where, detail)


def kind_id(name):
    k = ctypes.c_int32()
    check(lib().vj_hash_kind_from_name(name.encode(), ctypes.byref(k)), "vj_hash_kind_from_name(%r)" % name)
    return k.value


def hash_opts(kind="fnv", hash_bits=64, chunk_bits=64, word_mode=WORDS_REFERENCE):
    return HashOpts(kind_id(kind) if isinstance(kind, str) else int(kind), hash_bits, chunk_bits, word_mode)


def table(cols, rowids=None, key_signed=False, n_rows=None):
    """vj_table over device tensors. `cols` is a list of 1-D int32/uint32/int64 torch tensors."""
    t = Table()
    t.n_rows = int(cols[0].numel()) if n_rows is None else int(n_rows)
    t.n_cols = len(cols)
    t.key_width = cols[0].element_size()
    t.key_signed = 1 if key_signed else 0
    for i, c in enumerate(cols):
        assert c.is_contiguous() and c.element_size() == t.key_width
        t.cols[i] = c.data_ptr()
    t.rowids = None if rowids is None else rowids.data_ptr()
    t._keep = (cols, rowids)
    return t


def stream_ptr(stream=None):
    import torch
    s = torch.cuda.current_stream() if stream is None else stream
    return ctypes.c_void_p(s.cuda_stream)


def hash_table(cols, kind="fnv", hash_bits=64, chunk_bits=64, word_mode=WORDS_REFERENCE, key_signed=False,
               out=None):
    import torch
    t = table(cols, key_signed=key_signed)
    dt = torch.int32 if hash_bits == 32 else torch.int64
    if out is None:
        out = torch.empty(t.n_rows, dtype=dt, device=cols[0].device)
    o = hash_opts(kind, hash_bits, chunk_bits, word_mode)
    check(lib().vj_hash(ctypes.byref(t), ctypes.byref(o), ctypes.c_void_p(out.data_ptr()), stream_ptr()), "vj_hash")
    return out


def join_opts(kind="fnv", hash_bits=64, chunk_bits=64, word_mode=WORDS_REFERENCE, passes=-1, bits1=0, bits2=0,
              one_based=False, out_capacity_hint=0, tile=0, table=0, layout=0):
    return JoinOpts(hash_opts(kind, hash_bits, chunk_bits, word_mode), passes, bits1, bits2, 1 if one_based else 0,
                    out_capacity_hint, tile, table, layout, 0)


class Context:
    """Owns a vj_ctx (one per device)."""

    def __init__(self, device=0):
        self.ptr = ctypes.c_void_p()
        check(lib().vj_ctx_create(device, ctypes.byref(self.ptr)), "vj_ctx_create")

    def close(self):
        if self.ptr:
            lib().vj_ctx_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_budget(self, nbytes):
        check(lib().vj_ctx_set_budget(self.ptr, int(nbytes)), "vj_ctx_set_budget", self.ptr)

    def set_profiling(self, on):
        check(lib().vj_ctx_set_profiling(self.ptr, 1 if on else 0), "vj_ctx_set_profiling", self.ptr)

    def memory(self):
        b = ctypes.c_uint64()
        check(lib().vj_ctx_memory(self.ptr, ctypes.byref(b)), "vj_ctx_memory", self.ptr)
        return b.value

    def last_error(self):
        m = lib().vj_last_error(self.ptr)
        return m.decode() if m else ""

    def join_into(self, build, probe, opts, out_r, out_s, capacity=None):
        """Returns (status, count). Does not raise on VJ_ERR_CAPACITY."""
        cnt = ctypes.c_uint64()
        cap = int(out_r.numel()) if capacity is None else int(capacity)
        rc = lib().vj_join_into(self.ptr, ctypes.byref(build), ctypes.byref(probe), ctypes.byref(opts),
                                ctypes.c_void_p(out_r.data_ptr() if out_r is not None else 0),
                                ctypes.c_void_p(out_s.data_ptr() if out_s is not None else 0), cap,
                                ctypes.byref(cnt), stream_ptr())
        return rc, cnt.value

    def join_into_noalloc(self, build, probe, opts, out_r, out_s, capacity=None):
        """vj_join_into_noalloc (fixed per-device scratch, nothing allocated); returns (status, count)."""
        cnt = ctypes.c_uint64()
        cap = int(out_r.numel()) if capacity is None else int(capacity)
        rc = lib().vj_join_into_noalloc(self.ptr, ctypes.byref(build), ctypes.byref(probe), ctypes.byref(opts),
                                        ctypes.c_void_p(out_r.data_ptr()), ctypes.c_void_p(out_s.data_ptr()), cap,
                                        ctypes.byref(cnt), stream_ptr())
        return rc, cnt.value

    def join(self, build, probe, opts):
        """vj_join; returns (build_rowids, probe_rowids) as torch int32 tensors (copies; frees the vj_pairs)."""
        import torch
        p = Pairs()
        check(lib().vj_join(self.ptr, ctypes.byref(build), ctypes.byref(probe), ctypes.byref(opts), ctypes.byref(p),
                            stream_ptr()), "vj_join", self.ptr)
        try:
            n = int(p.count)
            r = torch.empty(n, dtype=torch.int32, device="cuda")
            s = torch.empty(n, dtype=torch.int32, device="cuda")
            if n:
                _copy_dev(r, p.build_rowids, n * 4)
                _copy_dev(s, p.probe_rowids, n * 4)
            torch.cuda.synchronize()
        finally:
            lib().vj_pairs_free(ctypes.byref(p))
        return r, s

    def join_rows(self, r_values, r_pks, s_values, s_pks, opts):
        """vj_join_rows on reference-layout tables (device int64 tensors: values [n, cols] row-major, pks [n]).
        Returns (rs_values [m, 2], rs_pks [m]) as torch int64 tensors (copies; frees the callee's table)."""
        import torch

        def rows(v, pk):
            t = Rows()
            t.n_rows = int(pk.numel())
            t.n_cols = int(v.shape[1])
            t.column_values = v.data_ptr() if t.n_rows else None
            t.primary_keys = pk.data_ptr() if t.n_rows else None
            return t
        r, s, rs = rows(r_values, r_pks), rows(s_values, s_pks), Rows()
        check(lib().vj_join_rows(self.ptr, ctypes.byref(r), ctypes.byref(s), ctypes.byref(opts), ctypes.byref(rs),
                                 stream_ptr()), "vj_join_rows", self.ptr)
        try:
            m = int(rs.n_rows)
            assert m == 0 or rs.n_cols == 2
            vals = torch.empty((m, 2), dtype=torch.int64, device="cuda")
            pks = torch.empty(m, dtype=torch.int64, device="cuda")
            if m:
                _copy_dev(vals, rs.column_values, m * 16)
                _copy_dev(pks, rs.primary_keys, m * 8)
            torch.cuda.synchronize()
        finally:
            lib().vj_rows_free(ctypes.byref(rs))
        return vals, pks

    def summary(self):
        s = JoinSummary()
        check(lib().vj_join_summary_get(self.ptr, ctypes.byref(s)), "vj_join_summary_get", self.ptr)
        return {f: getattr(s, f) for f, _ in JoinSummary._fields_}

    def partition(self, tab, hopts, shift, bits, out_tab, hist, offsets):
        return lib().vj_partition(self.ptr, ctypes.byref(tab), ctypes.byref(hopts), shift, bits,
                                  ctypes.byref(out_tab), ctypes.c_void_p(hist.data_ptr()),
                                  ctypes.c_void_p(offsets.data_ptr()), stream_ptr())

    def tune(self, build, probe, opts, reps=3):
        """vj_tune: returns ((passes, bits1, bits2, tile, table, layout), median ms) of the fastest plan, kept
        for the size class (tile / table / layout: TILE_* / TABLE_* / LAYOUT_*)."""
        best = JoinOpts()
        ms = ctypes.c_double()
        check(lib().vj_tune(self.ptr, ctypes.byref(build), ctypes.byref(probe), ctypes.byref(opts), int(reps),
                            ctypes.byref(best), ctypes.byref(ms), stream_ptr()), "vj_tune", self.ptr)
        return (best.passes, best.bits1, best.bits2, best.tile, best.table, best.layout), ms.value

    def tune_clear(self):
        check(lib().vj_tune_clear(self.ptr), "vj_tune_clear", self.ptr)

    def build(self, table, opts):
        """vj_build: a resident, partitioned build side (Built)."""
        h = ctypes.c_void_p()
        check(lib().vj_build(self.ptr, ctypes.byref(table), ctypes.byref(opts), ctypes.byref(h), stream_ptr()),
              "vj_build", self.ptr)
        return Built(h)

    def probe_into(self, built, probe, opts, out_r, out_s, capacity=None):
        """vj_probe_into; returns (status, count). Does not raise on VJ_ERR_CAPACITY."""
        cnt = ctypes.c_uint64()
        cap = int(out_r.numel()) if capacity is None else int(capacity)
        rc = lib().vj_probe_into(self.ptr, built.ptr, ctypes.byref(probe), ctypes.byref(opts),
                                 ctypes.c_void_p(out_r.data_ptr()), ctypes.c_void_p(out_s.data_ptr()), cap,
                                 ctypes.byref(cnt), stream_ptr())
        if rc not in (VJ_OK, STATUS_CODE["VJ_ERR_CAPACITY"]):
            check(rc, "vj_probe_into", self.ptr)
        return rc, cnt.value

    def join_batch(self, builds, probes, opts, outs_r, outs_s, caps, use_graph=False):
        n = len(builds)
        B = (Table * n)(*builds)
        S = (Table * n)(*probes)
        R_ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs_r])
        S_ptrs = (ctypes.c_void_p * n)(*[o.data_ptr() for o in outs_s])
        C = (ctypes.c_uint64 * n)(*[int(c) for c in caps])
        counts = (ctypes.c_uint64 * n)()
        rc = lib().vj_join_batch(self.ptr, n, B, S, ctypes.byref(opts), R_ptrs, S_ptrs, C, counts,
                                 1 if use_graph else 0, stream_ptr())
        return rc, [counts[i] for i in range(n)]


class Built:
    """A resident build side (vj_built)."""

    def __init__(self, ptr):
        self.ptr = ptr

    def nbytes(self):
        b = ctypes.c_uint64()
        check(lib().vj_built_bytes(self.ptr, ctypes.byref(b)), "vj_built_bytes")
        return b.value

    def free(self):
        if self.ptr:
            lib().vj_built_free(self.ptr)
            self.ptr = ctypes.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Server:
    """The serving engine (vj_server): submit requests, wait for responses.
    Tables and output tensors of a request must stay alive until it completes;
    the server keeps references to them until then."""

    def __init__(self, device=0, budget=0, opts=None, max_batch=64):
        self.ptr = ctypes.c_void_p()
        o = opts if opts is not None else join_opts()
        check(lib().vj_server_create(device, int(budget), ctypes.byref(o), int(max_batch), ctypes.byref(self.ptr)),
              "vj_server_create")
        self._live = {}

    def submit(self, probe, out_r, out_s, build=None, table=0, keep_build=False, capacity=None):
        req = Request()
        req.build = ctypes.cast(ctypes.pointer(build), ctypes.c_void_p) if build is not None else None
        req.table = int(table)
        req.probe = ctypes.cast(ctypes.pointer(probe), ctypes.c_void_p)
        req.out_build = out_r.data_ptr()
        req.out_probe = out_s.data_ptr()
        req.capacity = int(out_r.numel()) if capacity is None else int(capacity)
        req.keep_build = 1 if keep_build else 0
        t = ctypes.c_uint64()
        check(lib().vj_server_submit(self.ptr, ctypes.byref(req), ctypes.byref(t)), "vj_server_submit")
        self._live[t.value] = (build, probe, out_r, out_s)
        return t.value

    def wait(self, ticket):
        r = Response()
        check(lib().vj_server_wait(self.ptr, int(ticket), ctypes.byref(r)), "vj_server_wait")
        self._live.pop(ticket, None)
        return {"status": r.status, "status_name": STATUS.get(r.status, str(r.status)), "count": r.count,
                "table": r.table, "batch_size": r.batch_size, "t_submit_ns": r.t_submit_ns,
                "t_launch_ns": r.t_launch_ns, "t_done_ns": r.t_done_ns}

    def drop_table(self, table):
        check(lib().vj_server_drop_table(self.ptr, int(table)), "vj_server_drop_table")

    def stats(self):
        st = ServerStats()
        check(lib().vj_server_stats_get(self.ptr, ctypes.byref(st)), "vj_server_stats_get")
        return {n: getattr(st, n) for n, _ in ServerStats._fields_}

    def close(self):
        if self.ptr:
            lib().vj_server_destroy(self.ptr)
            self.ptr = ctypes.c_void_p()
            self._live.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class _DevArray:
    """Zero-copy view of a raw device pointer for torch.as_tensor (__cuda_array_interface__)."""

    def __init__(self, ptr, n, typestr="<i4"):
        self.__cuda_array_interface__ = {"shape": (n,), "typestr": typestr, "data": (int(ptr), False),
                                         "version": 2, "strides": None}


def _copy_dev(dst_tensor, src_ptr, nbytes):
    """Byte copy of nbytes (a multiple of 4) from a raw device pointer into a contiguous tensor."""
    import torch
    n = nbytes // 4
    view = torch.as_tensor(_DevArray(src_ptr, n), device=dst_tensor.device)
    dst_tensor.view(-1).view(torch.int32)[:n].copy_(view)
