"""ctypes binding of libdnspir.so (the HIP engine behind include/dnspir.h).

The product path always runs through this library; there is no CPU fallback.
Importing works without a GPU (the client functions are host code); device
calls fail with DNSPIR_E_NOINIT when no GPU is present.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libdnspir.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "dnspir.h")

N = 2048
LANES = 8

ERRORS = {
    -1: "DNSPIR_E_SHORT_QUERY",
    -2: "DNSPIR_E_CONJ",
    -3: "DNSPIR_E_COUNT",
    -4: "DNSPIR_E_SHAPE",
    -5: "DNSPIR_E_ARG",
    -6: "DNSPIR_E_ALIGN",
    -7: "DNSPIR_E_OOM",
    -8: "DNSPIR_E_HIP",
    -9: "DNSPIR_E_DECODE",
    -10: "DNSPIR_E_NOINIT",
    -11: "DNSPIR_E_NOKEY",
    -12: "DNSPIR_E_CLOSED",
    -13: "DNSPIR_E_IO",
    -14: "DNSPIR_E_FORMAT",
    -15: "DNSPIR_E_SCRATCH",
    -16: "DNSPIR_E_FRAME",
}
E_NOKEY = -11
QUERY_KEYLESS = 1


class DnspirError(RuntimeError):
    def __init__(self, code, where):
        self.code = code
        super().__init__(f"{where}: {ERRORS.get(code, code)} ({code})")


class Timing(ctypes.Structure):
    _fields_ = [("primary_us", ctypes.c_double), ("secondary_us", ctypes.c_double),
                ("total_us", ctypes.c_double), ("bytes_streamed", ctypes.c_uint64),
                ("k1_us", ctypes.c_double), ("k1_launches", ctypes.c_int), ("ref_bytes", ctypes.c_uint64)]


class ServerConfig(ctypes.Structure):
    _fields_ = [("slots", ctypes.c_int), ("max_batch", ctypes.c_int), ("scratch_budget", ctypes.c_size_t),
                ("key_cache_entries", ctypes.c_int), ("batch_window", ctypes.c_double)]


class RequestInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("db", ctypes.c_int), ("batch_size", ctypes.c_int),
                ("key_cached", ctypes.c_int), ("resent", ctypes.c_int), ("submit_us", ctypes.c_double),
                ("start_us", ctypes.c_double), ("done_us", ctypes.c_double)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("batches", ctypes.c_uint64),
                ("key_hits", ctypes.c_uint64), ("key_inserts", ctypes.c_uint64), ("key_evictions", ctypes.c_uint64),
                ("key_misses", ctypes.c_uint64), ("admission_waits", ctypes.c_uint64),
                ("scratch_in_use_peak", ctypes.c_size_t), ("inflight_peak", ctypes.c_int),
                ("pair_refused", ctypes.c_uint64), ("pair_retries", ctypes.c_uint64)]


class CtxStats(ctypes.Structure):
    _fields_ = [("pair_runs", ctypes.c_uint64), ("pair_refused", ctypes.c_uint64), ("pair_retries", ctypes.c_uint64),
                ("quad_runs", ctypes.c_uint64)]


class FrontendConfig(ctypes.Structure):
    _fields_ = [("server", ctypes.c_void_p), ("db", ctypes.c_int), ("num_entries", ctypes.c_size_t),
                ("zone_name", ctypes.c_char_p), ("hash_salt", ctypes.c_uint64), ("dump", ctypes.c_void_p),
                ("dump_len", ctypes.c_size_t)]


class FrontendStats(ctypes.Structure):
    _fields_ = [(k, ctypes.c_uint64) for k in ("connections", "frames_in", "frames_out", "bytes_in", "bytes_out",
                                                "queries", "infos", "dumps", "unavailable", "errors")]


class TuneEntry(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 48), ("column_tiles", ctypes.c_int), ("ms", ctypes.c_double),
                ("exact", ctypes.c_int), ("chosen", ctypes.c_int)]


class Arrival(ctypes.Structure):
    _fields_ = [("db", ctypes.c_int), ("client", ctypes.c_uint64), ("arrival_us", ctypes.c_double),
                ("query", ctypes.c_void_p), ("qlen", ctypes.c_size_t), ("keys", ctypes.c_void_p)]


vp = ctypes.c_void_p
sz = ctypes.c_size_t
ci = ctypes.c_int
u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u64p = ctypes.POINTER(ctypes.c_uint64)
ip = ctypes.POINTER(ctypes.c_int)

SIGNATURES = {
    "dnspir_init": (ci, [ci]),
    "dnspir_strerror": (ctypes.c_char_p, [ci]),
    "dnspir_last_error": (ctypes.c_char_p, []),
    "dnspir_required_bytes": (sz, [ci]),
    "dnspir_shard_create": (ci, [ci, ci, vp, sz, ctypes.POINTER(vp)]),
    "dnspir_shard_destroy": (ci, [vp]),
    "dnspir_shard_group_order": (ci, [vp]),
    "dnspir_shard_save": (ci, [vp, ctypes.c_char_p]),
    "dnspir_shard_file_info": (ci, [ctypes.c_char_p, ctypes.POINTER(ci), ctypes.POINTER(ci), ctypes.POINTER(sz)]),
    "dnspir_shard_load": (ci, [ctypes.c_char_p, vp, sz, ctypes.POINTER(vp)]),
    "dnspir_shard_set_db": (ci, [vp, u16p, sz]),
    "dnspir_shard_set_db_device": (ci, [vp, vp, sz]),
    "dnspir_shard_get_a": (ci, [vp, u64p]),
    "dnspir_shard_set_db_buckets": (ci, [vp, u8p, sz]),
    "dnspir_pack_bytes_into_coeffs": (ci, [u8p, sz, u16p]),
    "dnspir_frontend_create": (ci, [ctypes.POINTER(FrontendConfig), ctypes.POINTER(vp)]),
    "dnspir_frontend_handle_fd": (ci, [vp, ci]),
    "dnspir_frontend_listen": (ci, [vp, ctypes.c_char_p, ci, ctypes.POINTER(ci)]),
    "dnspir_frontend_get_stats": (ci, [vp, ctypes.POINTER(FrontendStats)]),
    "dnspir_frontend_destroy": (ci, [vp]),
    "dnspir_shard_perform_pir": (ci, [vp, u64p, u64p, u64p, u64p]),
    "dnspir_shard_perform_pir_split": (ci, [vp, u64p, u64p, u64p, u64p]),
    "dnspir_ctx_create": (ci, [vp, ctypes.POINTER(vp)]),
    "dnspir_ctx_destroy": (ci, [vp]),
    "dnspir_ctx_stream": (vp, [vp]),
    "dnspir_ctx_set_graphs": (ci, [vp, ci]),
    "dnspir_ctx_set_exchange_polls": (ci, [vp, ci]),
    "dnspir_ctx_set_chain_cus": (ci, [vp, ci]),
    "dnspir_ctx_set_k1_mode": (ci, [vp, ci]),
    "dnspir_ctx_get_stats": (ci, [vp, ctypes.POINTER(CtxStats)]),
    "dnspir_double_pir": (ci, [vp, vp, sz, vp, sz, u64p, u64p, u64p, u64p, ctypes.POINTER(Timing)]),
    "dnspir_process_queries": (ci, [vp, vp, sz, vp, sz, u8p, sz, u8p, sz, u8p, u8p, ctypes.POINTER(Timing)]),
    "dnspir_process_queries_device": (ci, [vp, vp, sz, vp, sz, vp, sz, vp, sz, vp, vp, ctypes.POINTER(Timing)]),
    "dnspir_process_queries_noalloc": (ci, [vp, sz, vp, sz, vp, sz, vp, sz, vp, vp, ctypes.POINTER(Timing)]),
    "dnspir_process_batch": (ci, [vp, vp, sz, u8p, sz, ci, u8p, ctypes.POINTER(Timing)]),
    "dnspir_process_batch_device": (ci, [vp, vp, sz, vp, sz, ci, vp, ctypes.POINTER(Timing)]),
    "dnspir_database_shape": (ci, [sz, ip, ip]),
    "dnspir_num_databases": (sz, [sz]),
    "dnspir_choose_index_group_secondary": (ci, [sz]),
    "dnspir_query_bytes": (sz, [ci, ci]),
    "dnspir_reply_bytes": (sz, []),
    "dnspir_rng_new": (vp, [u8p]),
    "dnspir_rng_free": (None, [vp]),
    "dnspir_prepare_query": (ci, [vp, sz, sz, ci, ci, u8p, ctypes.POINTER(sz), u8p]),
    "dnspir_process_reply": (ci, [u8p, sz, u8p, u16p]),
    "dnspir_client_keys": (ci, [vp, u8p, u8p]),
    "dnspir_client_query": (ci, [vp, u8p, sz, sz, ci, ci, u8p, ctypes.POINTER(sz)]),
    "dnspir_key_bytes": (sz, []),
    "dnspir_autotune_batch": (ci, [vp, vp, sz, ci, ctypes.POINTER(TuneEntry), ci, ip]),
    "dnspir_probe_mfma_rate": (ci, [vp, ci, ci, ctypes.POINTER(ctypes.c_double)]),
    "dnspir_probe_read_rate": (ci, [vp, vp, sz, ci, ctypes.POINTER(ctypes.c_double)]),
    "dnspir_batch_scratch_bytes": (sz, [ci, sz, ci]),
    "dnspir_server_create": (ci, [ctypes.POINTER(ServerConfig), ctypes.POINTER(vp)]),
    "dnspir_server_add_database": (ci, [vp, vp, sz, ip]),
    "dnspir_server_max_batch": (ci, [vp, ci]),
    "dnspir_server_submit": (ci, [vp, ci, ctypes.c_uint64, u8p, sz, ci, ctypes.POINTER(ctypes.c_uint64)]),
    "dnspir_server_wait": (ci, [vp, ctypes.c_uint64, u8p, ctypes.POINTER(RequestInfo)]),
    "dnspir_server_replay": (ci, [vp, ctypes.POINTER(Arrival), sz, u8p, ctypes.POINTER(RequestInfo)]),
    "dnspir_server_get_stats": (ci, [vp, ctypes.POINTER(ServerStats)]),
    "dnspir_server_destroy": (ci, [vp]),
}

_lib = None


def lib():
    """Load libdnspir.so; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build the HIP engine first (__graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, where):
    if rc != 0:
        if rc == -8:
            where = f"{where} [{lib().dnspir_last_error().decode()}]"
        raise DnspirError(rc, where)
    return rc
