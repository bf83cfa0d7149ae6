"""ctypes binding of the CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, the smoke check and the CPU-baseline leg of bench.py import this
module. It never serves a product call: the product path is the HIP library
under dnspir_amd/.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "build", "liborc.so")

P1 = 67104769
P2 = 33550337
Q = P1 * P2
T = 1025
LANES = 8

_lib = None

u8p = ctypes.POINTER(ctypes.c_uint8)
u16p = ctypes.POINTER(ctypes.c_uint16)
u32p = ctypes.POINTER(ctypes.c_uint32)
u64p = ctypes.POINTER(ctypes.c_uint64)
i64p = ctypes.POINTER(ctypes.c_int64)
vp = ctypes.c_void_p
sz = ctypes.c_size_t


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        sig = {
            "orc_chacha_block": (None, [u32p, u32p, ctypes.c_int]),
            "orc_rng_new": (vp, [u8p]),
            "orc_rng_free": (None, [vp]),
            "orc_rng_next_u32": (ctypes.c_uint32, [vp]),
            "orc_rng_next_u64": (ctypes.c_uint64, [vp]),
            "orc_expand": (None, [ctypes.c_int, u8p, u64p]),
            "orc_seed": (None, [ctypes.c_int, u8p]),
            "orc_gen_sk": (None, [ctypes.c_int, vp, ctypes.c_int, i64p]),
            "orc_enc_sym_b": (None, [ctypes.c_int, u8p, vp, i64p, u16p, ctypes.c_uint64, ctypes.c_double, u64p]),
            "orc_gen_gk_b": (None, [ctypes.c_int, vp, i64p, ctypes.c_int, ctypes.c_double, u64p]),
            "orc_dec": (None, [ctypes.c_int, i64p, u64p, u64p, ctypes.c_uint64, u16p]),
            "orc_poly_mul_sparse": (None, [ctypes.c_int, u64p, i64p, u64p]),
            "orc_mod_switch_encode": (None, [ctypes.c_int, ctypes.c_int, u64p, u64p]),
            "orc_mod_switch_decode": (None, [ctypes.c_int, ctypes.c_int, u64p, u64p]),
            "orc_ntt": (ctypes.c_uint64, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, u64p]),
            "orc_shard_create": (vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_shard_free": (None, [vp]),
            "orc_shard_set_transformed_entries": (ctypes.c_int, [vp, u16p, sz]),
            "orc_shard_preprocess": (ctypes.c_int, [vp]),
            "orc_shard_set_db": (ctypes.c_int, [vp, u16p, sz]),
            "orc_shard_get_a": (None, [vp, u64p]),
            "orc_shard_perform_pir_split": (ctypes.c_int, [vp, u64p, u64p, u64p, u64p]),
            "orc_shard_perform_pir": (ctypes.c_int, [vp, u64p, u64p, u64p, u64p]),
            "orc_shard_export_te_p1": (ctypes.c_uint64, [vp, u32p]),
            "orc_shard_export_row_coeffs": (None, [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, i64p]),
            "orc_shard_loop_len": (ctypes.c_int, [vp]),
            "orc_shard_requires_conj": (ctypes.c_int, [vp]),
            "orc_double_pir": (ctypes.c_int, [vp, sz, vp, sz, u64p, u64p, ctypes.c_int, u64p, u64p,
                                              ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
            "orc_choose_index_group_secondary": (ctypes.c_int, [ctypes.c_int, sz]),
            "orc_condense": (sz, [ctypes.c_int, u64p, sz, u8p]),
            "orc_uncondense": (ctypes.c_int, [ctypes.c_int, sz, u8p, sz, ctypes.POINTER(sz), u64p]),
            "orc_shard_k1_operands": (ctypes.c_int, [vp, u64p, u64p, u64p, u32p, u32p, ctypes.POINTER(ctypes.c_int)]),
            "orc_shard_k1_finish": (ctypes.c_int, [vp, u64p, u64p]),
            "orc_cpu_k1": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(u32p), ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                          u32p, ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int, u64p]),
            "orc_cpu_isa": (ctypes.c_char_p, []),
            "orc_cpu_reduce_every": (ctypes.c_int, []),
            "orc_cpu_force_isa": (None, [ctypes.c_int]),
            "orc_pack_bytes_into_coeffs": (None, [u8p, u16p]),
            "orc_unpack_coeffs_into_bytes": (None, [u16p, u8p]),
            "orc_database_shape": (None, [sz, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
            "orc_num_databases": (sz, [sz]),
            "orc_query_bytes": (sz, [ctypes.c_int, ctypes.c_int]),
            "orc_process_queries": (ctypes.c_int, [vp, sz, vp, sz, u8p, sz, u8p, sz, u8p, u8p,
                                                   ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double)]),
            "orc_prepare_query": (ctypes.c_int, [vp, sz, sz, ctypes.c_int, ctypes.c_int, u8p, ctypes.POINTER(sz), u8p]),
            "orc_process_reply": (ctypes.c_int, [u8p, sz, u8p, u16p]),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_set_reuse_repeated": (None, [ctypes.c_int]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def ptr(a, t):
    return a.ctypes.data_as(t)


def set_threads(n):
    lib().orc_set_threads(int(n))


def set_reuse_repeated(on):
    """Repeated shards of a fleet reuse their first occurrence's primary
    result (a test speed-up; identical output)."""
    lib().orc_set_reuse_repeated(1 if on else 0)


def chacha_block(inp, rounds):
    inp = np.ascontiguousarray(inp, dtype=np.uint32)
    out = np.zeros(16, np.uint32)
    lib().orc_chacha_block(ptr(inp, u32p), ptr(out, u32p), rounds)
    return out


class Rng:
    def __init__(self, seed):
        s = np.frombuffer(bytes(seed), dtype=np.uint8).copy()
        assert s.size == 32
        self.h = lib().orc_rng_new(ptr(s, u8p))

    def next_u32(self):
        return lib().orc_rng_next_u32(self.h)

    def next_u64(self):
        return lib().orc_rng_next_u64(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_rng_free(self.h)
            self.h = None


def seed(i):
    out = np.zeros(32, np.uint8)
    lib().orc_seed(i, ptr(out, u8p))
    return bytes(out)


def expand(logN, seed_bytes):
    s = np.frombuffer(bytes(seed_bytes), np.uint8).copy()
    out = np.zeros(1 << logN, np.uint64)
    lib().orc_expand(logN, ptr(s, u8p), ptr(out, u64p))
    return out


def ntt(logN, modulus, x, inverse=False, root_pow=1):
    x = np.ascontiguousarray(x, dtype=np.uint64).copy()
    z = lib().orc_ntt(logN, modulus, root_pow, 1 if inverse else 0, ptr(x, u64p))
    return x, z


def gen_sk(logN, rng, hwt=128):
    sk = np.zeros(1 << logN, np.int64)
    lib().orc_gen_sk(logN, rng.h, hwt, ptr(sk, i64p))
    return sk


def enc_sym_b(logN, seed_a, rng, sk, m, t=T, sigma=3.2):
    s = np.frombuffer(bytes(seed_a), np.uint8).copy()
    m = np.ascontiguousarray(m, dtype=np.uint16)
    b = np.zeros(1 << logN, np.uint64)
    lib().orc_enc_sym_b(logN, ptr(s, u8p), rng.h, ptr(sk, i64p), ptr(m, u16p), t, sigma, ptr(b, u64p))
    return b


def gen_gk_b(logN, rng, sk, g=5, sigma=3.2):
    gk = np.zeros((2, 1 << logN), np.uint64)
    lib().orc_gen_gk_b(logN, rng.h, ptr(sk, i64p), g, sigma, ptr(gk, u64p))
    return gk


def dec(logN, sk, a, b, t=T):
    a = np.ascontiguousarray(a, np.uint64)
    b = np.ascontiguousarray(b, np.uint64)
    m = np.zeros(1 << logN, np.uint16)
    lib().orc_dec(logN, ptr(sk, i64p), ptr(a, u64p), ptr(b, u64p), t, ptr(m, u16p))
    return m


def mod_switch_encode(logN, bits, x):
    x = np.ascontiguousarray(x, np.uint64)
    out = np.zeros(1 << logN, np.uint64)
    lib().orc_mod_switch_encode(logN, bits, ptr(x, u64p), ptr(out, u64p))
    return out


def mod_switch_decode(logN, bits, x):
    x = np.ascontiguousarray(x, np.uint64)
    out = np.zeros(1 << logN, np.uint64)
    lib().orc_mod_switch_decode(logN, bits, ptr(x, u64p), ptr(out, u64p))
    return out


class Shard:
    """PIRDatabase restatement (base_pir.rs)."""

    def __init__(self, logN, group_order, used_seeds=0, root_pow=1):
        self.logN = logN
        self.N = 1 << logN
        self.l = group_order
        self.h = lib().orc_shard_create(logN, group_order, used_seeds, root_pow)
        if not self.h:
            raise ValueError("invalid index group")
        self.L = lib().orc_shard_loop_len(self.h)
        self.conj = bool(lib().orc_shard_requires_conj(self.h))

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_shard_free(self.h)
            self.h = None

    def set_transformed_entries(self, entries):
        e = np.ascontiguousarray(entries, np.uint16)
        rc = lib().orc_shard_set_transformed_entries(self.h, ptr(e, u16p), e.shape[0])
        if rc:
            raise ValueError(f"set_transformed_entries failed: {rc}")

    def preprocess(self):
        lib().orc_shard_preprocess(self.h)

    def set_db(self, entries):
        e = np.ascontiguousarray(entries, np.uint16)
        rc = lib().orc_shard_set_db(self.h, ptr(e, u16p), e.shape[0])
        if rc:
            raise ValueError(f"set_db failed: {rc}")

    def get_a(self):
        out = np.zeros((LANES, self.N), np.uint64)
        lib().orc_shard_get_a(self.h, ptr(out, u64p))
        return out

    def perform_pir(self, qb, qbc, gkb):
        qb = np.ascontiguousarray(qb, np.uint64)
        gkb = np.ascontiguousarray(gkb, np.uint64)
        qbc_p = ptr(np.ascontiguousarray(qbc, np.uint64), u64p) if qbc is not None else None
        out = np.zeros((LANES, self.N), np.uint64)
        rc = lib().orc_shard_perform_pir(self.h, ptr(qb, u64p), qbc_p, ptr(gkb, u64p), ptr(out, u64p))
        if rc:
            raise ValueError(f"perform_pir failed: {rc}")
        return out

    def perform_pir_split(self, qb, qbc, gkb):
        qb = np.ascontiguousarray(qb, np.uint64)
        gkb = np.ascontiguousarray(gkb, np.uint64)
        qbc_p = ptr(np.ascontiguousarray(qbc, np.uint64), u64p) if qbc is not None else None
        out = np.zeros((LANES, self.N), np.uint64)
        rc = lib().orc_shard_perform_pir_split(self.h, ptr(qb, u64p), qbc_p, ptr(gkb, u64p), ptr(out, u64p))
        if rc:
            raise ValueError(f"perform_pir_split failed: {rc}")
        return out

    def export_te_p1(self):
        rows = self.L * (2 if self.conj else 1)
        out = np.zeros((rows, LANES, self.N), np.uint32)
        z = lib().orc_shard_export_te_p1(self.h, ptr(out, u32p))
        return out, z

    def row_coeffs(self, array, stream, row, lane):
        out = np.zeros(self.N, np.int64)
        lib().orc_shard_export_row_coeffs(self.h, array, stream, row, lane, ptr(out, i64p))
        return out


def _fleet(shards):
    arr = (ctypes.c_void_p * max(1, len(shards)))(*[s.h for s in shards])
    return arr


def double_pir(fleet0, fleet1, q0, q1, l_s):
    N = fleet0[0].N
    f0 = _fleet(fleet0)
    f1 = _fleet(fleet1) if fleet1 else None
    qs = []
    for q in (q0, q1):
        qs.append(np.ascontiguousarray(q, np.uint64) if q is not None else None)
    outs = [np.zeros((3, 2, N), np.uint64) if q is not None else None for q in qs]
    pu = ctypes.c_double()
    su = ctypes.c_double()
    rc = lib().orc_double_pir(f0, len(fleet0), f1, len(fleet1) if fleet1 else 0,
                              ptr(qs[0], u64p) if qs[0] is not None else None,
                              ptr(qs[1], u64p) if qs[1] is not None else None,
                              l_s,
                              ptr(outs[0], u64p) if outs[0] is not None else None,
                              ptr(outs[1], u64p) if outs[1] is not None else None,
                              ctypes.byref(pu), ctypes.byref(su))
    if rc:
        raise ValueError(f"double_pir failed: {rc}")
    return outs, pu.value, su.value


def choose_index_group_secondary(num_shards, logN=11):
    return lib().orc_choose_index_group_secondary(logN, num_shards)


def condense(bits, data):
    d = np.ascontiguousarray(data, np.uint64)
    out = np.zeros(d.size * bits // 8 + 8, np.uint8)
    n = lib().orc_condense(bits, ptr(d, u64p), d.size, ptr(out, u8p))
    return bytes(out[:n])


def uncondense(bits, count, data):
    b = np.frombuffer(bytes(data), np.uint8).copy() if len(data) else np.zeros(1, np.uint8)
    out = np.zeros(count, np.uint64)
    used = sz()
    rc = lib().orc_uncondense(bits, count, ptr(b, u8p), len(data), ctypes.byref(used), ptr(out, u64p))
    if rc:
        raise ValueError("not enough data")
    return out, used.value


def database_shape(n):
    lp = ctypes.c_int()
    ls = ctypes.c_int()
    lib().orc_database_shape(n, ctypes.byref(lp), ctypes.byref(ls))
    return lp.value, ls.value


def num_databases(n):
    return lib().orc_num_databases(n)


def query_bytes(lp, ls):
    return lib().orc_query_bytes(lp, ls)


REPLY_BYTES = 3 * (2048 * 17 // 8 + 2048 * 12 // 8)


def process_queries(fleet0, fleet1, query0, query1):
    f0 = _fleet(fleet0)
    f1 = _fleet(fleet1) if fleet1 else None
    bufs = []
    for q in (query0, query1):
        bufs.append(np.frombuffer(bytes(q), np.uint8).copy() if q is not None else None)
    reps = [np.zeros(REPLY_BYTES, np.uint8) if q is not None else None for q in (query0, query1)]
    pu = ctypes.c_double()
    su = ctypes.c_double()
    rc = lib().orc_process_queries(f0, len(fleet0), f1, len(fleet1) if fleet1 else 0,
                                   ptr(bufs[0], u8p) if bufs[0] is not None else None, len(query0) if query0 is not None else 0,
                                   ptr(bufs[1], u8p) if bufs[1] is not None else None, len(query1) if query1 is not None else 0,
                                   ptr(reps[0], u8p) if reps[0] is not None else None,
                                   ptr(reps[1], u8p) if reps[1] is not None else None,
                                   ctypes.byref(pu), ctypes.byref(su))
    if rc:
        raise ValueError(f"process_queries failed: {rc}")
    return [bytes(r) if r is not None else None for r in reps], pu.value, su.value


def prepare_query(rng, primary_idx, secondary_idx, lp, ls):
    out = np.zeros(80000, np.uint8)
    n = sz()
    sk = np.zeros(32, np.uint8)
    rc = lib().orc_prepare_query(rng.h, primary_idx, secondary_idx, lp, ls, ptr(out, u8p), ctypes.byref(n), ptr(sk, u8p))
    if rc:
        raise ValueError(f"prepare_query failed: {rc}")
    return bytes(out[:n.value]), bytes(sk)


def process_reply(reply, sk_seed):
    r = np.frombuffer(bytes(reply), np.uint8).copy()
    s = np.frombuffer(bytes(sk_seed), np.uint8).copy()
    out = np.zeros(2048, np.uint16)
    rc = lib().orc_process_reply(ptr(r, u8p), len(reply), ptr(s, u8p), ptr(out, u16p))
    if rc:
        raise ValueError(f"process_reply failed: {rc}")
    return out


def pack_bytes_into_coeffs(buckets):
    """dns/bucket.rs:289-303: [n][2560] bucket bytes -> [n][2048] coefficients"""
    b = np.ascontiguousarray(buckets, np.uint8).reshape(-1, 2560)
    out = np.zeros((b.shape[0], 2048), np.uint16)
    for i in range(b.shape[0]):
        lib().orc_pack_bytes_into_coeffs(ptr(b[i], u8p), ptr(out[i], u16p))
    return out


def unpack_coeffs_into_bytes(coeffs):
    """dns/bucket.rs:271-283: [n][2048] coefficients -> [n][2560] bytes"""
    c = np.ascontiguousarray(coeffs, np.uint16).reshape(-1, 2048)
    out = np.zeros((c.shape[0], 2560), np.uint8)
    for i in range(c.shape[0]):
        lib().orc_unpack_coeffs_into_bytes(ptr(c[i], u16p), ptr(out[i], u8p))
    return out


class RefK1:
    """main_pir_loop in the reference's own form (oracle/cpu_baseline.c), for
    the CPU baseline: a shard's arrays and a query's vectors exported once in
    the reference layout, then the timed loop."""

    def __init__(self, shard, qb, qbc, gkb):
        """qb, qbc: [8][N] (qbc None without a conjugate), gkb: [8][2][N]; query k on lane k"""
        self.shard = shard
        N, L, nst = shard.N, shard.L, (2 if shard.conj else 1)
        self.dims = (N, L, nst)
        self.arrays = np.zeros(2 * nst * 3 * L * N * LANES, np.uint32)
        self.vecs = np.zeros(2 * nst * 3 * N * LANES, np.uint32)
        self.gal = np.zeros(N, np.int32)
        qb = np.ascontiguousarray(qb, np.uint64)
        gkb = np.ascontiguousarray(gkb, np.uint64)
        qc = ptr(np.ascontiguousarray(qbc, np.uint64), u64p) if qbc is not None else None
        rc = lib().orc_shard_k1_operands(shard.h, ptr(qb, u64p), qc, ptr(gkb, u64p), ptr(self.arrays, u32p),
                                         ptr(self.vecs, u32p), self.gal.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
        if rc:
            raise ValueError(f"k1_operands failed: {rc}")

    @staticmethod
    def run(parts, threads=1, window=0, keep=True):
        """the loops of every shard in `parts` (RefK1 objects of one query) ->
        sums [n][2][N][8] (None if keep is False)"""
        N, L, nst = parts[0].dims
        arr = (u32p * len(parts))(*[ptr(p.arrays, u32p) for p in parts])
        sums = np.zeros((len(parts), 2, N, LANES), np.uint64) if keep else None
        rc = lib().orc_cpu_k1(len(parts), arr, N, L, nst, ptr(parts[0].vecs, u32p),
                              parts[0].gal.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), threads, window,
                              ptr(sums, u64p) if keep else None)
        if rc:
            raise MemoryError("orc_cpu_k1")
        return sums

    def finish(self, sums):
        """sums [2][N][8] -> perform_pir_split's output [8][N] mod q"""
        s = np.ascontiguousarray(sums, np.uint64)
        out = np.zeros((LANES, self.dims[0]), np.uint64)
        lib().orc_shard_k1_finish(self.shard.h, ptr(s, u64p), ptr(out, u64p))
        return out


def cpu_isa():
    return lib().orc_cpu_isa().decode()


def cpu_reduce_every():
    return lib().orc_cpu_reduce_every()


def cpu_force_isa(isa):
    lib().orc_cpu_force_isa(isa)
