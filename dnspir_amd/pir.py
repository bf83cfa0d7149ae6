"""Host-side mirror of the `dnspir-pir` crate API over the HIP engine.

Names and argument meanings follow crates/pir/src/{base_pir,double_pir,
pir_wrapper}.rs: `PIRDatabase` (create / set_db / get_a / perform_pir /
perform_pir_split), `perform_batched_double_pir`, `process_queries`,
`prepare_query` / `process_reply`, `get_database_shape`, `num_databases`.
Errors the reference reports by panicking raise `DnspirError`.

Device memory for shards is caller-owned, as in the reference (the crate
borrows a slab, base_pir.rs:280-300); here a torch uint8 tensor on the GPU
holds it.
"""
import ctypes
import os

import numpy as np

from ._lib import (LANES, N, QUERY_KEYLESS, Arrival, CtxStats, DnspirError, FrontendConfig, FrontendStats, RequestInfo,
                   ServerConfig, ServerStats, Timing, TuneEntry, check, lib)

BUCKET_BYTES = 2560

_ptr = lambda a, t: a.ctypes.data_as(t)  # noqa: E731
u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u64p = ctypes.POINTER(ctypes.c_uint64)

_initialised = None


def init(device=0):
    """dnspir_init: builds the device tables (idempotent)."""
    global _initialised
    if _initialised != device:
        import torch  # device memory and streams only
        torch.cuda.init()
        check(lib().dnspir_init(device), "dnspir_init")
        _initialised = device
    return True


def required_bytes(group_order):
    return lib().dnspir_required_bytes(group_order)


def shard_file_info(path):
    """(group_order, used_seeds, payload bytes) from a shard file's header."""
    l, seeds, nb = ctypes.c_int(), ctypes.c_int(), ctypes.c_size_t()
    check(lib().dnspir_shard_file_info(os.fsencode(path), ctypes.byref(l), ctypes.byref(seeds), ctypes.byref(nb)),
          "dnspir_shard_file_info")
    return l.value, seeds.value, nb.value


class PIRDatabase:
    """One preprocessed shard (base_pir.rs:225-1247) resident in HBM."""

    def __init__(self, group_order=N, used_seeds=0, device=0, memory=None, _load_path=None):
        import torch
        init(device)
        self.group_order = group_order
        nbytes = required_bytes(group_order)
        if nbytes == 0:
            raise DnspirError(-4, "PIRDatabase")
        if memory is None:
            memory = torch.empty(nbytes + 256, dtype=torch.uint8, device=f"cuda:{device}")
        self.memory = memory  # keeps the slab alive
        base = memory.data_ptr()
        aligned = (base + 255) & ~255
        # the bytes the caller's memory actually holds past the aligned start
        avail = max(0, memory.numel() * memory.element_size() - (aligned - base))
        h = ctypes.c_void_p()
        if _load_path is None:
            check(lib().dnspir_shard_create(group_order, used_seeds, ctypes.c_void_p(aligned), avail, ctypes.byref(h)),
                  "dnspir_shard_create")
        else:
            check(lib().dnspir_shard_load(os.fsencode(_load_path), ctypes.c_void_p(aligned), avail, ctypes.byref(h)),
                  "dnspir_shard_load")
        self.h = h

    def save(self, path):
        """Writes the preprocessed shard's device state to `path` (checkpoint)."""
        check(lib().dnspir_shard_save(self.h, os.fsencode(path)), "dnspir_shard_save")

    @classmethod
    def load(cls, path, device=0, memory=None):
        """A shard resumed from a file written by save(), without preprocessing."""
        l, seeds, _ = shard_file_info(path)
        return cls(l, seeds, device, memory, _load_path=path)

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                lib().dnspir_shard_destroy(h)
            except Exception:  # interpreter shutdown: the library is already gone
                pass
            self.h = None

    @property
    def requires_conj(self):
        return self.group_order == N

    def set_db(self, entries):
        e = np.ascontiguousarray(entries, dtype=np.uint16)
        if e.ndim != 2 or e.shape[1] != N:
            raise DnspirError(-5, "set_db: entries must be [n][2048]")
        check(lib().dnspir_shard_set_db(self.h, _ptr(e, u16p), e.shape[0]), "set_db")

    def set_db_device(self, entries_dev):
        """entries_dev: a [n][2048] 16-bit torch tensor on this GPU."""
        if entries_dev.dim() != 2 or entries_dev.shape[1] != N or entries_dev.element_size() != 2:
            raise DnspirError(-5, "set_db_device: entries must be a [n][2048] 16-bit tensor")
        import torch
        e = entries_dev.contiguous()
        # the library works on its own streams: finish the producer's work first
        torch.cuda.current_stream().synchronize()
        check(lib().dnspir_shard_set_db_device(self.h, ctypes.c_void_p(e.data_ptr()), e.shape[0]), "set_db_device")

    def set_db_buckets(self, buckets):
        """set_db from up to 8*l DNS buckets of 2560 bytes ([n][2560] uint8),
        packed into coefficients on the device (dns/bucket.rs:289-303); the
        remaining entries of the shard are zero."""
        b = np.ascontiguousarray(buckets, dtype=np.uint8).reshape(-1, BUCKET_BYTES)
        check(lib().dnspir_shard_set_db_buckets(self.h, _ptr(b, u8p), b.shape[0]), "set_db_buckets")

    def get_a(self):
        out = np.zeros((LANES, N), np.uint64)
        check(lib().dnspir_shard_get_a(self.h, _ptr(out, u64p)), "get_a")
        return out

    def perform_pir(self, qry_b, qry_conj_b, gk_b):
        qb = np.ascontiguousarray(qry_b, np.uint64)
        gk = np.ascontiguousarray(gk_b, np.uint64)
        qc = _ptr(np.ascontiguousarray(qry_conj_b, np.uint64), u64p) if qry_conj_b is not None else None
        out = np.zeros((LANES, N), np.uint64)
        check(lib().dnspir_shard_perform_pir(self.h, _ptr(qb, u64p), qc, _ptr(gk, u64p), _ptr(out, u64p)), "perform_pir")
        return out

    def perform_pir_split(self, qry_b, qry_conj_b, gk_b):
        qb = np.ascontiguousarray(qry_b, np.uint64)
        gk = np.ascontiguousarray(gk_b, np.uint64)
        qc = _ptr(np.ascontiguousarray(qry_conj_b, np.uint64), u64p) if qry_conj_b is not None else None
        out = np.zeros((LANES, N), np.uint64)
        check(lib().dnspir_shard_perform_pir_split(self.h, _ptr(qb, u64p), qc, _ptr(gk, u64p), _ptr(out, u64p)),
              "perform_pir_split")
        return out


class Context:
    """A query context: per-stream secondary scratch (double_pir.rs:43-56)."""

    def __init__(self, stream=None):
        init()
        h = ctypes.c_void_p()
        check(lib().dnspir_ctx_create(ctypes.c_void_p(stream) if stream else None, ctypes.byref(h)), "dnspir_ctx_create")
        self.h = h

    @property
    def stream(self):
        return lib().dnspir_ctx_stream(self.h)

    def set_graphs(self, on):
        """Batched calls replay a captured HIP graph (default on); returns
        True if capture failed on this context (eager launches)."""
        rc = lib().dnspir_ctx_set_graphs(self.h, 1 if on else 0)
        if rc < 0:
            raise DnspirError(rc, "ctx_set_graphs")
        return rc == 1

    def set_exchange_polls(self, max_polls):
        """Bound of each partner wait of the multi-CU secondary chain (0: give
        up at once; the call then runs its secondary phase again on one CU
        per chain)."""
        check(lib().dnspir_ctx_set_exchange_polls(self.h, int(max_polls)), "ctx_set_exchange_polls")

    def set_chain_cus(self, cus):
        """Most CUs per secondary chain: 0 (default) four up to 16 chains, two
        up to 128, else one; or 1, 2, 4. Same bytes."""
        check(lib().dnspir_ctx_set_chain_cus(self.h, int(cus)), "ctx_set_chain_cus")

    def set_k1_mode(self, mode):
        """Single-query inner product: 3 streams all three shard arrays, 2
        streams TE and IPGK1 with per-query weights in place of IPGK0, 0
        (default) picks 2 for fleets of 8 or more shards. Same bytes."""
        check(lib().dnspir_ctx_set_k1_mode(self.h, int(mode)), "ctx_set_k1_mode")

    def stats(self):
        """counters of this context's secondary phases: multi-CU runs (and
        the four-CU ones among them), runs on one CU because the device's
        multi-CU budget was in use, retries after a partner wait timed out"""
        st = CtxStats()
        check(lib().dnspir_ctx_get_stats(self.h, ctypes.byref(st)), "ctx_get_stats")
        return {k: getattr(st, k) for k, _ in CtxStats._fields_}

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                lib().dnspir_ctx_destroy(h)
            except Exception:  # interpreter shutdown: the library is already gone
                pass
            self.h = None


def _fleet(shards):
    if not shards:
        return None, 0
    arr = (ctypes.c_void_p * len(shards))(*[s.h.value for s in shards])
    return arr, len(shards)


def perform_batched_double_pir(fleet0, fleet1, q0, q1, ctx=None):
    """double_pir.rs:126-228. q0/q1: [6][N] query blocks or None; returns
    ([3][2][N] or None per slot, Timing)."""
    init()
    f0, l0 = _fleet(fleet0)
    f1, l1 = _fleet(fleet1)
    qs = [np.ascontiguousarray(q, np.uint64) if q is not None else None for q in (q0, q1)]
    outs = [np.zeros((3, 2, N), np.uint64) if q is not None else None for q in qs]
    tm = Timing()
    check(lib().dnspir_double_pir(ctx.h if ctx else None, f0, l0, f1, l1,
                                  *[_ptr(q, u64p) if q is not None else None for q in qs],
                                  *[_ptr(o, u64p) if o is not None else None for o in outs],
                                  ctypes.byref(tm)), "double_pir")
    return outs, tm


def process_queries(fleet0, fleet1, query0, query1=None, ctx=None):
    """pir_wrapper.rs:349-409 (host bytes in, host bytes out)."""
    init()
    f0, l0 = _fleet(fleet0)
    f1, l1 = _fleet(fleet1)
    rb = lib().dnspir_reply_bytes()
    bufs = [np.frombuffer(bytes(q), np.uint8).copy() if q is not None else None for q in (query0, query1)]
    reps = [np.zeros(rb, np.uint8) if q is not None else None for q in (query0, query1)]
    tm = Timing()
    check(lib().dnspir_process_queries(ctx.h if ctx else None, f0, l0, f1, l1,
                                       _ptr(bufs[0], u8p) if bufs[0] is not None else None, len(bufs[0]) if bufs[0] is not None else 0,
                                       _ptr(bufs[1], u8p) if bufs[1] is not None else None, len(bufs[1]) if bufs[1] is not None else 0,
                                       _ptr(reps[0], u8p) if reps[0] is not None else None,
                                       _ptr(reps[1], u8p) if reps[1] is not None else None,
                                       ctypes.byref(tm)), "process_queries")
    return [bytes(r) if r is not None else None for r in reps], tm


def process_queries_device(fleet0, fleet1, query0_dev, query1_dev, reply0_dev, reply1_dev, ctx=None):
    """Same with query / reply bytes in device memory (torch uint8 tensors)."""
    import torch
    init()
    f0, l0 = _fleet(fleet0)
    f1, l1 = _fleet(fleet1)
    tm = Timing()
    # query bytes may come from torch's stream; the engine runs on the context's stream
    torch.cuda.current_stream().synchronize()
    dp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    check(lib().dnspir_process_queries_device(ctx.h if ctx else None, f0, l0, f1, l1,
                                              dp(query0_dev), query0_dev.numel() if query0_dev is not None else 0,
                                              dp(query1_dev), query1_dev.numel() if query1_dev is not None else 0,
                                              dp(reply0_dev), dp(reply1_dev), ctypes.byref(tm)),
          "process_queries_device")
    return tm


def process_queries_noalloc(fleet0, fleet1, query0_dev, query1_dev, reply0_dev, reply1_dev):
    """process_queries_device with all device scratch in the library's fixed
    per-GPU array (dnspir_process_queries_noalloc): allocates nothing."""
    import torch
    init()
    f0, l0 = _fleet(fleet0)
    f1, l1 = _fleet(fleet1)
    tm = Timing()
    torch.cuda.current_stream().synchronize()
    dp = lambda t: ctypes.c_void_p(t.data_ptr()) if t is not None else None  # noqa: E731
    check(lib().dnspir_process_queries_noalloc(f0, l0, f1, l1,
                                               dp(query0_dev), query0_dev.numel() if query0_dev is not None else 0,
                                               dp(query1_dev), query1_dev.numel() if query1_dev is not None else 0,
                                               dp(reply0_dev), dp(reply1_dev), ctypes.byref(tm)),
          "process_queries_noalloc")
    return tm


def process_batch(fleet, queries, ctx=None):
    """B independent queries against one fleet in one pass (host bytes);
    returns (list of reply bytes, Timing)."""
    init()
    f, n = _fleet(fleet)
    qlen = len(queries[0])
    if any(len(q) != qlen for q in queries):
        raise DnspirError(-5, "process_batch: all queries of a fleet have the same size")
    buf = np.frombuffer(b"".join(bytes(q) for q in queries), np.uint8).copy()
    rb = reply_bytes()
    out = np.zeros(len(queries) * rb, np.uint8)
    tm = Timing()
    check(lib().dnspir_process_batch(ctx.h if ctx else None, f, n, _ptr(buf, u8p), qlen, len(queries),
                                     _ptr(out, u8p), ctypes.byref(tm)), "process_batch")
    return [bytes(out[i * rb:(i + 1) * rb]) for i in range(len(queries))], tm


def process_batch_device(fleet, queries_dev, qlen, nq, replies_dev, ctx=None):
    """Batched queries with packed query / reply bytes in device memory."""
    import torch
    init()
    f, n = _fleet(fleet)
    tm = Timing()
    torch.cuda.current_stream().synchronize()
    check(lib().dnspir_process_batch_device(ctx.h if ctx else None, f, n, ctypes.c_void_p(queries_dev.data_ptr()), qlen, nq,
                                            ctypes.c_void_p(replies_dev.data_ptr()), ctypes.byref(tm)),
          "process_batch_device")
    return tm


def autotune_batch(fleet, nq, ctx=None):
    """dnspir_autotune_batch: times the batched inner-product candidates for
    batches of nq queries on this fleet, keeps the fastest exact one; returns
    [{name, column_tiles, ms, exact, chosen}]."""
    init()
    f, n = _fleet(fleet)
    ents = (TuneEntry * 32)()
    cnt = ctypes.c_int()
    check(lib().dnspir_autotune_batch(ctx.h if ctx else None, f, n, nq, ents, 32, ctypes.byref(cnt)), "autotune_batch")
    return [{"name": e.name.decode(), "column_tiles": e.column_tiles, "ms": round(e.ms, 4), "exact": bool(e.exact),
             "chosen": bool(e.chosen)} for e in ents[:cnt.value]]


def probe_mfma_rate(zero_operands=False, reps=5, ctx=None):
    """dnspir_probe_mfma_rate: int8 TOPS of the batched inner product's MFMA
    mix on register operands (hashed digits, or all zero)."""
    init()
    t = ctypes.c_double()
    check(lib().dnspir_probe_mfma_rate(ctx.h if ctx else None, 1 if zero_operands else 0, reps, ctypes.byref(t)),
          "probe_mfma_rate")
    return t.value


def probe_read_rate(fleet, reps=3, ctx=None):
    """dnspir_probe_read_rate: GB/s of a plain streaming read of the fleet's
    shard slabs (the same 16-byte loads as the inner products)."""
    init()
    f, n = _fleet(fleet)
    t = ctypes.c_double()
    check(lib().dnspir_probe_read_rate(ctx.h if ctx else None, f, n, reps, ctypes.byref(t)), "probe_read_rate")
    return t.value


def get_database_shape(num_entries):
    """pir_wrapper.rs:219-231; DnspirError(DNSPIR_E_SHAPE) where the reference
    asserts (0 entries, more than N^2)."""
    lp, ls = ctypes.c_int(), ctypes.c_int()
    check(lib().dnspir_database_shape(num_entries, ctypes.byref(lp), ctypes.byref(ls)), "get_database_shape")
    return lp.value, ls.value


def num_databases(num_entries):
    n = lib().dnspir_num_databases(num_entries)
    if n == 0:
        raise DnspirError(-4, "num_databases")
    return n


def choose_index_group_secondary(num_shards):
    """pir_wrapper.rs:166-175; DnspirError(DNSPIR_E_SHAPE) above N/8 shards."""
    rc = lib().dnspir_choose_index_group_secondary(num_shards)
    if rc < 0:
        raise DnspirError(rc, "choose_index_group_secondary")
    return rc


def query_bytes(lp, ls):
    return lib().dnspir_query_bytes(lp, ls)


def reply_bytes():
    return lib().dnspir_reply_bytes()


class Rng:
    """StdRng stand-in (ChaCha12) for the client side."""

    def __init__(self, seed):
        s = np.frombuffer(bytes(seed), np.uint8).copy()
        assert s.size == 32
        self.h = lib().dnspir_rng_new(_ptr(s, u8p))

    def __del__(self):
        h = getattr(self, "h", None)
        if h:
            try:
                lib().dnspir_rng_free(h)
            except Exception:  # interpreter shutdown: the library is already gone
                pass
            self.h = None


def prepare_query(rng, primary_idx, secondary_idx, l_primary, l_secondary):
    """prepare_query_general (pir_wrapper.rs:276-317) -> (query bytes, sk seed)."""
    out = np.zeros(query_bytes(l_primary, l_secondary) + 16, np.uint8)
    n = ctypes.c_size_t()
    sk = np.zeros(32, np.uint8)
    check(lib().dnspir_prepare_query(rng.h, primary_idx, secondary_idx, l_primary, l_secondary, _ptr(out, u8p),
                                     ctypes.byref(n), _ptr(sk, u8p)), "prepare_query")
    return bytes(out[:n.value]), bytes(sk)


def key_bytes():
    return lib().dnspir_key_bytes()


def client_keys(rng):
    """Key half of prepare_query: -> (key bytes, sk seed)."""
    sk = np.zeros(32, np.uint8)
    out = np.zeros(key_bytes(), np.uint8)
    check(lib().dnspir_client_keys(rng.h, _ptr(sk, u8p), _ptr(out, u8p)), "client_keys")
    return bytes(out), bytes(sk)


def client_query(rng, sk_seed, primary_idx, secondary_idx, l_primary, l_secondary):
    """Query half of prepare_query (no key prefix) for a client's long-lived key."""
    sk = np.frombuffer(bytes(sk_seed), np.uint8).copy()
    out = np.zeros(query_bytes(l_primary, l_secondary) + 16, np.uint8)
    n = ctypes.c_size_t()
    check(lib().dnspir_client_query(rng.h, _ptr(sk, u8p), primary_idx, secondary_idx, l_primary, l_secondary,
                                    _ptr(out, u8p), ctypes.byref(n)), "client_query")
    return bytes(out[:n.value])


def batch_scratch_bytes(l_primary, num_shards, nq):
    return lib().dnspir_batch_scratch_bytes(l_primary, num_shards, nq)


class Server:
    """Continuous-batching request loop over resident databases
    (dnspir_server_*): slots, HBM scratch budget, device Galois-key cache."""

    def __init__(self, slots=3, max_batch=64, scratch_budget=8 << 30, key_cache_entries=256, batch_window=0.0):
        init()
        cfg = ServerConfig(slots, max_batch, scratch_budget, key_cache_entries, batch_window)
        h = ctypes.c_void_p()
        check(lib().dnspir_server_create(ctypes.byref(cfg), ctypes.byref(h)), "dnspir_server_create")
        self.h = h
        self._fleets = []

    def add_database(self, fleet):
        arr, n = _fleet(fleet)
        db = ctypes.c_int()
        check(lib().dnspir_server_add_database(self.h, arr, n, ctypes.byref(db)), "server_add_database")
        self._fleets.append(list(fleet))  # the shards must outlive the server
        return db.value

    def max_batch(self, db):
        rc = lib().dnspir_server_max_batch(self.h, db)
        if rc < 0:
            raise DnspirError(rc, "server_max_batch")
        return rc

    def submit(self, db, client, query, keyless=False):
        """Queues one query; returns a ticket (raises DnspirError, e.g.
        DNSPIR_E_NOKEY for a keyless query whose keys are not cached)."""
        q = np.frombuffer(bytes(query), np.uint8).copy()
        t = ctypes.c_uint64()
        check(lib().dnspir_server_submit(self.h, db, client, _ptr(q, u8p), q.size, QUERY_KEYLESS if keyless else 0,
                                         ctypes.byref(t)), "server_submit")
        return t.value

    def wait(self, ticket):
        """-> (reply bytes, RequestInfo)"""
        out = np.zeros(reply_bytes(), np.uint8)
        info = RequestInfo()
        check(lib().dnspir_server_wait(self.h, ticket, _ptr(out, u8p), ctypes.byref(info)), "server_wait")
        return bytes(out), info

    def replay(self, arrivals):
        """arrivals: list of (db, client, arrival_us, keyless query bytes, key
        bytes). Submits each at its time (open loop), waits for all; returns
        (replies as one uint8 array [n][reply_bytes], list of RequestInfo)."""
        n = len(arrivals)
        arr = (Arrival * n)()
        bufs = {}  # one C copy per distinct bytes object

        def cbuf(b):
            if b is None:
                return None
            k = id(b)
            if k not in bufs:
                bufs[k] = (b, ctypes.create_string_buffer(bytes(b), len(b)))
            return ctypes.cast(bufs[k][1], ctypes.c_void_p)

        for i, (db, client, t_us, q, keys) in enumerate(arrivals):
            arr[i] = Arrival(db, client, t_us, cbuf(q), len(q), cbuf(keys))
        replies = np.zeros((n, reply_bytes()), np.uint8)
        infos = (RequestInfo * n)()
        lib().dnspir_server_replay(self.h, arr, n, _ptr(replies, u8p), infos)
        return replies, list(infos)

    def stats(self):
        st = ServerStats()
        check(lib().dnspir_server_get_stats(self.h, ctypes.byref(st)), "server_get_stats")
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def close(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib().dnspir_server_destroy(h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library is already gone
            pass


def pack_bytes_into_coeffs(buckets):
    """dns/bucket.rs:289-303 on the host: [n][2560] bytes -> [n][2048] uint16"""
    b = np.ascontiguousarray(buckets, dtype=np.uint8).reshape(-1, BUCKET_BYTES)
    out = np.zeros((b.shape[0], N), np.uint16)
    check(lib().dnspir_pack_bytes_into_coeffs(_ptr(b, u8p), b.shape[0], _ptr(out, u16p)), "pack_bytes_into_coeffs")
    return out


class Frontend:
    """The reference server's TCP protocol (dns/protocol.rs, server.rs:747-807)
    in front of a serving loop database: INFO / QUERY / DUMP frames."""

    def __init__(self, server=None, db=0, num_entries=0, zone_name="", hash_salt=0, dump=None):
        self._dump = None if dump is None else ctypes.create_string_buffer(bytes(dump), len(dump))
        self._zone = zone_name.encode()
        cfg = FrontendConfig(server.h if server is not None else None, db, num_entries, self._zone, hash_salt,
                             ctypes.cast(self._dump, ctypes.c_void_p) if self._dump is not None else None,
                             len(dump) if dump is not None else 0)
        h = ctypes.c_void_p()
        check(lib().dnspir_frontend_create(ctypes.byref(cfg), ctypes.byref(h)), "dnspir_frontend_create")
        self.h = h
        self.server = server  # must outlive the front end

    def handle_fd(self, fd):
        """serves one connection on fd until the peer closes; returns the status"""
        return lib().dnspir_frontend_handle_fd(self.h, fd)

    def listen(self, host="127.0.0.1", port=0):
        bound = ctypes.c_int()
        check(lib().dnspir_frontend_listen(self.h, host.encode(), port, ctypes.byref(bound)), "frontend_listen")
        return bound.value

    def stats(self):
        st = FrontendStats()
        check(lib().dnspir_frontend_get_stats(self.h, ctypes.byref(st)), "frontend_get_stats")
        return {k: getattr(st, k) for k, _ in FrontendStats._fields_}

    def close(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            lib().dnspir_frontend_destroy(h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:  # interpreter shutdown: the library is already gone
            pass


def process_reply(reply, sk_seed):
    """pir_wrapper.rs:420-440 -> N coefficients mod 1025."""
    r = np.frombuffer(bytes(reply), np.uint8).copy()
    s = np.frombuffer(bytes(sk_seed), np.uint8).copy()
    out = np.zeros(N, np.uint16)
    check(lib().dnspir_process_reply(_ptr(r, u8p), len(reply), _ptr(s, u8p), _ptr(out, u16p)), "process_reply")
    return out
