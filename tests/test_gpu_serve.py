"""GPU serving loop (dnspir_server_*): continuous batching over three
resident databases of different shapes. Every reply is byte-identical to the
same query answered alone (process_queries) and decrypts to the requested
entry; keyless queries use the device key cache (LRU); admission keeps the
in-flight scratch within the budget."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
N = 2048


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pir():
    from dnspir_amd import pir as p
    p.init(0)
    return p


@pytest.fixture(scope="module")
def dbs(pir):
    """three fleets: 2 half shards (l_s 16), 3 half shards (l_s 32),
    2 full shards (l_s 16, conjugated primary queries)"""
    out = []
    for d, (l, nsh) in enumerate([(1024, 2), (1024, 3), (2048, 2)]):
        _say(f"[serve] preprocessing db {d}: {nsh} shards l={l}")
        fleet, ents = [], []
        for k in range(nsh):
            e = np.random.default_rng(1000 * d + k).integers(0, 1024, (8 * l, N)).astype(np.uint16)
            s = pir.PIRDatabase(l)
            s.set_db(e)
            fleet.append(s)
            ents.append(e)
        out.append((fleet, ents, l, pir.choose_index_group_secondary(nsh)))
    return out


class Client:
    def __init__(self, pir, cid):
        self.pir = pir
        self.cid = cid
        self.rng = pir.Rng(bytes([cid % 251, cid // 251 + 1] + [3] * 30))
        self.keys, self.sk = pir.client_keys(self.rng)

    def query(self, db, seed):
        fleet, ents, l, ls = db
        r = np.random.default_rng(seed)
        pi, si = int(r.integers(0, 8 * l)), int(r.integers(0, len(fleet)))
        return pi, si, self.pir.client_query(self.rng, self.sk, pi, si, l, ls)


def _alone(pir, db, full_query):
    (rep, _), _ = pir.process_queries(db[0], None, full_query, None)
    return rep


def test_serve_mix_matches_alone(pir, dbs):
    srv = pir.Server(slots=3, max_batch=8, scratch_budget=4 << 30, key_cache_entries=64)
    ids = [srv.add_database(d[0]) for d in dbs]
    clients = [Client(pir, c) for c in range(12)]
    jobs = []
    for k in range(36):
        c = clients[k % len(clients)]
        d = (k * 7) % 3
        pi, si, q = c.query(dbs[d], k)
        full = c.keys + q
        jobs.append((c, d, pi, si, full, srv.submit(ids[d], c.cid, full)))
    _say("[serve] 36 requests submitted")
    sizes = []
    for j, (c, d, pi, si, full, t) in enumerate(jobs):
        rep, info = srv.wait(t)
        assert info.status == 0 and info.db == ids[d]
        sizes.append(info.batch_size)
        assert np.array_equal(pir.process_reply(rep, c.sk), dbs[d][1][si][pi]), j
        if j % 6 == 0:
            assert rep == _alone(pir, dbs[d], full), j
    st = srv.stats()
    _say(f"[serve] stats {st} batch sizes {sorted(set(sizes))}")
    assert st["completed"] == 36 and st["failed"] == 0
    assert st["scratch_in_use_peak"] <= 4 << 30
    srv.close()


def test_keyless_queries_use_the_key_cache(pir, dbs):
    from dnspir_amd import DnspirError
    srv = pir.Server(slots=2, max_batch=4, scratch_budget=2 << 30, key_cache_entries=2)
    ids = [srv.add_database(d[0]) for d in dbs]
    a, b, c = Client(pir, 101), Client(pir, 102), Client(pir, 103)
    # no keys yet: keyless fails
    _, _, q = a.query(dbs[0], 1)
    with pytest.raises(DnspirError) as ei:
        srv.submit(ids[0], a.cid, q, keyless=True)
    assert ei.value.code == -11
    # full query stores the keys; a keyless one right after waits for them
    pi, si, q1 = a.query(dbs[2], 2)
    t1 = srv.submit(ids[2], a.cid, a.keys + q1)
    pi2, si2, q2 = a.query(dbs[1], 3)
    t2 = srv.submit(ids[1], a.cid, q2, keyless=True)
    r1, i1 = srv.wait(t1)
    r2, i2 = srv.wait(t2)
    assert i1.key_cached == 0 and i2.key_cached == 1
    assert np.array_equal(pir.process_reply(r1, a.sk), dbs[2][1][si][pi])
    assert np.array_equal(pir.process_reply(r2, a.sk), dbs[1][1][si2][pi2])
    assert r2 == _alone(pir, dbs[1], a.keys + q2)
    # two more clients evict a (capacity 2, least recently used first)
    for cl in (b, c):
        _, _, q = cl.query(dbs[0], 4)
        srv.wait(srv.submit(ids[0], cl.cid, cl.keys + q))
    _, _, q = a.query(dbs[0], 5)
    with pytest.raises(DnspirError):
        srv.submit(ids[0], a.cid, q, keyless=True)
    pi, si, q = c.query(dbs[2], 6)
    r, info = srv.wait(srv.submit(ids[2], c.cid, q, keyless=True))
    assert info.key_cached == 1
    assert np.array_equal(pir.process_reply(r, c.sk), dbs[2][1][si][pi])
    st = srv.stats()
    assert st["key_evictions"] >= 1 and st["key_misses"] == 2 and st["key_hits"] == 2
    srv.close()


def test_admission_keeps_scratch_within_budget(pir, dbs):
    need = [pir.batch_scratch_bytes(d[2], len(d[0]), 1) for d in dbs]
    budget = max(need) * 2 + (1 << 20)  # room for about two single-query batches
    srv = pir.Server(slots=4, max_batch=16, scratch_budget=budget, key_cache_entries=0)
    ids = [srv.add_database(d[0]) for d in dbs]
    assert all(1 <= srv.max_batch(i) < 16 for i in ids[2:])
    clients = [Client(pir, 200 + c) for c in range(6)]
    jobs = []
    for k in range(18):
        c = clients[k % 6]
        d = k % 3
        pi, si, q = c.query(dbs[d], 100 + k)
        jobs.append((c, d, pi, si, srv.submit(ids[d], c.cid, c.keys + q)))
    for c, d, pi, si, t in jobs:
        rep, info = srv.wait(t)
        assert np.array_equal(pir.process_reply(rep, c.sk), dbs[d][1][si][pi])
    st = srv.stats()
    _say(f"[serve] budget {budget} stats {st}")
    assert st["scratch_in_use_peak"] <= budget
    assert st["completed"] == 18
    srv.close()


def test_replay_open_loop_poisson(pir, dbs):
    srv = pir.Server(slots=3, max_batch=16, scratch_budget=4 << 30, key_cache_entries=8)
    ids = [srv.add_database(d[0]) for d in dbs]
    clients = [Client(pir, 300 + c) for c in range(10)]
    r = np.random.default_rng(5)
    t = np.cumsum(r.exponential(1e6 / 200.0, 60))  # 200 requests/s
    arr, want = [], []
    for k in range(60):
        c = clients[int(r.integers(0, 10))]
        d = int(r.choice(3, p=[0.2, 0.3, 0.5]))
        pi, si, q = c.query(dbs[d], 500 + k)
        arr.append((ids[d], c.cid, float(t[k]), q, c.keys))
        want.append((c, d, pi, si))
    replies, infos = srv.replay(arr)
    for k, (c, d, pi, si) in enumerate(want):
        assert infos[k].status == 0, k
        assert np.array_equal(pir.process_reply(replies[k].tobytes(), c.sk), dbs[d][1][si][pi]), k
    resent = sum(i.resent for i in infos)
    assert 10 <= resent < 60  # first contact of each client, plus evictions
    lat = [(i.done_us - i.submit_us) / 1e3 for i in infos]
    _say(f"[serve] replay p50 {np.percentile(lat, 50):.2f} ms p99 {np.percentile(lat, 99):.2f} ms, "
         f"resent {resent}, stats {srv.stats()}")
    srv.close()


def test_batching_window_keeps_replies_exact(pir, dbs):
    """With a batching window the server holds a busy database's requests back
    to fill larger batches; every reply stays the query's own."""
    srv = pir.Server(slots=3, max_batch=16, scratch_budget=4 << 30, key_cache_entries=16, batch_window=2.0)
    ids = [srv.add_database(d[0]) for d in dbs]
    clients = [Client(pir, 400 + c) for c in range(8)]
    r = np.random.default_rng(9)
    arr, want = [], []
    t = 0.0
    for k in range(48):
        c = clients[k % 8]
        d = int(r.choice(3, p=[0.2, 0.3, 0.5]))
        pi, si, q = c.query(dbs[d], 700 + k)
        t += float(r.exponential(300.0))  # ~3300 requests/s: the GPU stays busy
        arr.append((ids[d], c.cid, t, q, c.keys))
        want.append((c, d, pi, si))
    replies, infos = srv.replay(arr)
    for k, (c, d, pi, si) in enumerate(want):
        assert infos[k].status == 0, k
        assert np.array_equal(pir.process_reply(replies[k].tobytes(), c.sk), dbs[d][1][si][pi]), k
    st = srv.stats()
    _say(f"[serve] window: {st['batches']} batches for 48 requests")
    assert st["completed"] == 48 and st["batches"] < 48
    srv.close()


def test_concurrent_two_cu_batches_share_the_device(pir):
    """Four slots each take a batch of 42 queries (126 secondary chains: the
    largest two-CU chain launch, 256 blocks) of four databases at once. Only
    as many two-CU chain blocks as the GPU has CUs may be in flight, so the
    batches that find the budget taken run their chains on one CU each; no
    request fails and every reply equals the query answered alone."""
    fleet, ents = [], []
    for k in range(4):
        e = np.random.default_rng(4200 + k).integers(0, 1024, (8 * 1024, N)).astype(np.uint16)
        s = pir.PIRDatabase(1024)
        s.set_db(e)
        fleet.append(s)
        ents.append(e)
    ls = pir.choose_index_group_secondary(4)
    fleets = [[fleet[(k + d) % 4] for k in range(4)] for d in range(4)]  # four databases over the same shards
    srv = pir.Server(slots=4, max_batch=42, scratch_budget=6 << 30, key_cache_entries=0)
    ids = [srv.add_database(f) for f in fleets]
    assert all(srv.max_batch(i) == 42 for i in ids)
    rng = pir.Rng(bytes([42] * 32))
    jobs = []
    for j in range(4 * 42):
        d = j % 4
        r = np.random.default_rng(j)
        pi, si = int(r.integers(0, 8 * 1024)), int(r.integers(0, 4))
        q, sk = pir.prepare_query(rng, pi, si, 1024, ls)
        jobs.append((d, pi, si, q, sk))
    _say("[serve] 168 pair-sized requests: submitting")
    tickets = [srv.submit(ids[d], 9000 + i, q) for i, (d, pi, si, q, sk) in enumerate(jobs)]
    for i, ((d, pi, si, q, sk), t) in enumerate(zip(jobs, tickets)):
        rep, info = srv.wait(t)
        assert info.status == 0
        want = ents[(si + d) % 4][pi]
        assert np.array_equal(pir.process_reply(rep, sk), want), i
        if i % 7 == 0:
            assert rep == _alone(pir, (fleets[d],), q), i
    st = srv.stats()
    _say(f"[serve] stats {st}")
    assert st["failed"] == 0 and st["completed"] == 168
    srv.close()
