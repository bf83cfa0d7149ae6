"""Byte parity at the benchmark configurations' real shapes: GPU replies equal
the CPU oracle's reply bytes for the same seeded database and queries, and
decrypt to the requested entry (pir_wrapper.rs:593-595).

Entries are 2560 seeded random bytes packed into 2048 ten-bit coefficients
(pack_bytes_into_coeffs, dns/bucket.rs:289-303). Fleets reuse a few resident
shards cyclically, as `--preprocessed-db-count` does (pir_wrapper.rs:578); the
oracle computes each distinct shard's primary result once per query.
"""
import sys

import numpy as np
import pytest

import pyoracle as orc

pytestmark = pytest.mark.gpu
N = 2048

# bench_wrapped_pir's literal (primary index, shard) list, pir_wrapper.rs:562
REF_INDICES = [(0, 0), (1, 0), (0, 1), (1, 1), (2, 2), (1, 2), (2, 0), (2, 1), (2, 2), (2742, 128), (923, 255),
               (1389, 129), (200, 200), (4000, 255), (3000, 0), (3000, 1)]


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pack(raw):
    """5 bytes -> 4 ten-bit coefficients, little endian (bucket.rs:289-303)."""
    b = raw.reshape(raw.shape[0], 512, 5).astype(np.uint64)
    v = b[..., 0] | (b[..., 1] << 8) | (b[..., 2] << 16) | (b[..., 3] << 24) | (b[..., 4] << 32)
    c = (v[..., None] >> np.array([0, 10, 20, 30], np.uint64)) & 1023
    return c.reshape(raw.shape[0], N).astype(np.uint16)


@pytest.fixture(scope="module")
def pir():
    from dnspir_amd import pir as p
    p.init(0)
    orc.set_reuse_repeated(True)
    yield p
    orc.set_reuse_repeated(False)


_shards = {}


def resident(pir, l, k):
    """Resident shard k of order l: GPU shard, oracle shard, entries."""
    key = (l, k)
    if key not in _shards:
        _say(f"[configs] preprocessing shard l={l} #{k} (GPU and oracle)")
        raw = np.random.default_rng(7000 + 31 * l + k).integers(0, 256, (8 * l, 2560), dtype=np.uint8)
        e = pack(raw)
        g = pir.PIRDatabase(l)
        g.set_db(e)
        o = orc.Shard(11, l)
        o.set_db(e)
        _shards[key] = (g, o, e)
    return _shards[key]


def fleet(pir, entries, resident_count):
    lp, ls = pir.get_database_shape(entries)
    n = pir.num_databases(entries)
    res = [resident(pir, lp, k) for k in range(resident_count)]
    return ([res[k % resident_count][0] for k in range(n)], [res[k % resident_count][1] for k in range(n)],
            [r[2] for r in res], lp, ls, n)


def _check(pir, fg, fo, ents, lp, ls, n, queries, oracle_every=1, tag=""):
    resident_count = len(ents)
    for t, (pi, si, q, sk) in enumerate(queries):
        reps, _ = pir.process_queries(fg, None, q, None)
        if t % oracle_every == 0:
            _say(f"[configs] {tag} query {t}: oracle")
            oreps, _, _ = orc.process_queries(fo, None, q, None)
            assert reps[0] == oreps[0], (tag, t)
        want = ents[si % resident_count][pi]
        assert np.array_equal(pir.process_reply(reps[0], sk), want), (tag, t)


def test_cpu_case_32768_entries(pir):
    """`bench --db-entries 32768 --preprocessed-db-count 2`: 4 half shards, 2
    resident; the reference's literal index list (rng seed 0) and 16 seeded
    indices with the client rng seeded 1."""
    fg, fo, ents, lp, ls, n = fleet(pir, 32768, 2)
    assert (lp, ls, n) == (1024, 32, 4)
    rng = pir.Rng(bytes([0] + [0] * 31))
    qs = []
    for i, j in REF_INDICES:
        q, sk = pir.prepare_query(rng, i, j % n, lp, ls)
        qs.append((i, j % n, q, sk))
    _check(pir, fg, fo, ents, lp, ls, n, qs, tag="32768 reference list")
    rng = pir.Rng(bytes([1] + [0] * 31))
    idx = np.random.default_rng(0)
    qs = []
    for _ in range(16):
        i, j = int(idx.integers(0, 8 * lp)), int(idx.integers(0, n))
        q, sk = pir.prepare_query(rng, i, j, lp, ls)
        qs.append((i, j, q, sk))
    _check(pir, fg, fo, ents, lp, ls, n, qs, tag="32768 seeded")


def test_one_full_shard_32_queries(pir):
    """One full-size shard (16384 entries): conjugated primary query, 32
    seeded indices."""
    g, o, e = resident(pir, 2048, 0)
    ls = pir.choose_index_group_secondary(1)
    rng = pir.Rng(bytes([2] + [0] * 31))
    idx = np.random.default_rng(2)
    qs = []
    for _ in range(32):
        i = int(idx.integers(0, 8 * 2048))
        q, sk = pir.prepare_query(rng, i, 0, 2048, ls)
        qs.append((i, 0, q, sk))
    _check(pir, [g], [o], [e], 2048, ls, 1, qs, tag="full shard")


def test_2p22_shape_256_full_shards(pir):
    """2^22 entries = 256 full shards (2 resident here), full-group secondary
    with conjugated queries on both levels; 8 seeded indices, 3 against the
    oracle."""
    fg, fo, ents, lp, ls, n = fleet(pir, 1 << 22, 2)
    assert (lp, ls, n) == (2048, 2048, 256)
    rng = pir.Rng(bytes([3] + [0] * 31))
    idx = np.random.default_rng(3)
    qs = []
    for t in range(8):
        i, j = int(idx.integers(0, 8 * lp)), int(idx.integers(0, n))
        if t == 7:
            i, j = 8 * lp - 1, n - 1  # last entry of the last shard
        q, sk = pir.prepare_query(rng, i, j, lp, ls)
        qs.append((i, j, q, sk))
    _check(pir, fg, fo, ents, lp, ls, n, qs, oracle_every=3, tag="2^22")


def test_batched_262144_shape(pir):
    """262144 entries = 32 half shards (2 resident): 64 clients with their own
    keys in one batch; every reply equals the query's single reply, three equal
    the oracle's, all decrypt."""
    fg, fo, ents, lp, ls, n = fleet(pir, 262144, 2)
    assert (lp, ls, n) == (1024, 256, 32)
    rng = pir.Rng(bytes([4] + [0] * 31))
    idx = np.random.default_rng(4)
    qs = []
    for _ in range(64):
        i, j = int(idx.integers(0, 8 * lp)), int(idx.integers(0, n))
        q, sk = pir.prepare_query(rng, i, j, lp, ls)
        qs.append((i, j, q, sk))
    _say("[configs] batch of 64 at 262144 entries")
    reps, _ = pir.process_batch(fg, [q for _, _, q, _ in qs])
    for t, (i, j, q, sk) in enumerate(qs):
        if t % 8 == 0:
            single, _ = pir.process_queries(fg, None, q, None)
            assert reps[t] == single[0], t
        if t in (0, 31, 63):
            oreps, _, _ = orc.process_queries(fo, None, q, None)
            assert reps[t] == oreps[0], t
        assert np.array_equal(pir.process_reply(reps[t], sk), ents[j % 2][i]), t


def test_serving_mix_three_databases(pir):
    """The serving mix's three database shapes (32768, 262144 and 2^22 entries,
    2 resident shards each): requests of 24 clients split 20/30/50, keyless
    after first contact; each reply equals the query answered alone."""
    dbs = [fleet(pir, e, 2) for e in (32768, 262144, 1 << 22)]
    srv = pir.Server(slots=3, max_batch=16, scratch_budget=16 << 30, key_cache_entries=16)
    ids = [srv.add_database(d[0]) for d in dbs]
    r = np.random.default_rng(5)
    clients = []
    for c in range(24):
        rng = pir.Rng(bytes([5, c] + [0] * 30))
        keys, sk = pir.client_keys(rng)
        clients.append((rng, keys, sk))
    arr, want = [], []
    t_us = 0.0
    for k in range(48):
        c = int(r.integers(0, 24))
        d = int(r.choice(3, p=[0.2, 0.3, 0.5]))
        fg, fo, ents, lp, ls, n = dbs[d]
        i, j = int(r.integers(0, 8 * lp)), int(r.integers(0, n))
        rng, keys, sk = clients[c]
        q = pir.client_query(rng, sk, i, j, lp, ls)
        t_us += float(r.exponential(2000.0))
        arr.append((ids[d], c, t_us, q, keys))
        want.append((d, i, j, keys + q, sk))
    _say("[configs] serving mix: 48 requests")
    replies, infos = srv.replay(arr)
    srv.close()
    for k, (d, i, j, full, sk) in enumerate(want):
        assert infos[k].status == 0, k
        rep = replies[k].tobytes()
        fg, fo, ents, lp, ls, n = dbs[d]
        assert np.array_equal(pir.process_reply(rep, sk), ents[j % 2][i]), k
        if k % 6 == 0:
            alone, _ = pir.process_queries(fg, None, full, None)
            assert rep == alone[0], k
