"""The DNS-over-PIR front end on the GPU: a database of 20000 buckets (three
half shards, the last one padded with zero entries) built from 2560-byte
buckets on the device, served through the serving loop behind the reference's
TCP protocol (dns/protocol.rs, server.rs:747-807). A client that reads INFO,
derives the shape from num_entries (get_database_shape) and sends QUERY frames
over loopback TCP gets reply frames byte-identical to process_queries' reply
and to the oracle's for the same query, which decrypt and unpack
(bucket.rs:271-283) to the requested bucket -- zero bytes in the padding."""
import socket
import struct
import sys
import threading

import numpy as np
import pytest

import pyoracle as orc

pytestmark = pytest.mark.gpu
N = 2048
ENTRIES = 20000


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def frame(payload):
    return struct.pack(">I", len(payload)) + payload


def recv_exact(s, n):
    out = b""
    while len(out) < n:
        c = s.recv(n - len(out))
        if not c:
            raise EOFError
        out += c
    return out


def read_frame(s):
    (n,) = struct.unpack(">I", recv_exact(s, 4))
    return recv_exact(s, n)


@pytest.fixture(scope="module")
def setup():
    from dnspir_amd import pir
    pir.init(0)
    lp, ls = pir.get_database_shape(ENTRIES)
    nsh = pir.num_databases(ENTRIES)
    assert (lp, ls, nsh) == (1024, 32, 3)
    buckets = np.random.default_rng(2560).integers(0, 256, (ENTRIES, 2560), dtype=np.uint8)
    fleet, ofleet = [], []
    for k in range(nsh):
        _say(f"[frontend] shard {k}: GPU from buckets, oracle from packed coefficients")
        chunk = buckets[k * 8 * lp:(k + 1) * 8 * lp]
        g = pir.PIRDatabase(lp)
        g.set_db_buckets(chunk)
        coeffs = np.zeros((8 * lp, N), np.uint16)
        coeffs[:len(chunk)] = orc.pack_bytes_into_coeffs(chunk)
        o = orc.Shard(11, lp)
        o.set_db(coeffs)
        assert np.array_equal(g.get_a(), o.get_a()), k
        fleet.append(g)
        ofleet.append(o)
    srv = pir.Server(slots=2, max_batch=8, scratch_budget=2 << 30, key_cache_entries=16)
    db = srv.add_database(fleet)
    fe = pir.Frontend(server=srv, db=db, num_entries=ENTRIES, zone_name="", hash_salt=0x5A17)
    port = fe.listen("127.0.0.1", 0)
    yield pir, buckets, fleet, ofleet, srv, fe, port
    fe.close()
    srv.close()


def _query(pir, rng, k):
    lp, ls = pir.get_database_shape(ENTRIES)
    return pir.prepare_query(rng, k % (8 * lp), k // (8 * lp), lp, ls)


def _expect(buckets, k):
    return buckets[k] if k < ENTRIES else np.zeros(2560, np.uint8)


def test_info_then_queries_over_tcp(setup):
    pir, buckets, fleet, ofleet, srv, fe, port = setup
    s = socket.create_connection(("127.0.0.1", port), timeout=60)
    s.sendall(frame(b"\x01"))
    p = read_frame(s)
    flags, num = p[0], struct.unpack(">I", p[1:5])[0]
    assert flags == 1 and num == ENTRIES  # PIR available, no dump
    rng = pir.Rng(bytes([3] * 32))
    for k in (0, 8191, 8192 + 17, ENTRIES - 1, ENTRIES + 5):
        q, sk = _query(pir, rng, k)
        _say(f"[frontend] QUERY for entry {k}")
        s.sendall(frame(b"\x02" + q))
        rep = read_frame(s)
        alone, _ = pir.process_queries(fleet, None, q, None)
        oreps, _, _ = orc.process_queries(ofleet, None, q, None)
        assert rep == alone[0] == oreps[0], k
        got = orc.unpack_coeffs_into_bytes(pir.process_reply(rep, sk))[0]
        assert np.array_equal(got, _expect(buckets, k)), k
    s.close()


def test_concurrent_connections_are_batched(setup):
    pir, buckets, fleet, ofleet, srv, fe, port = setup
    before = srv.stats()["batches"]
    results, errs = {}, []

    def client(cid):
        try:
            rng = pir.Rng(bytes([cid + 10] * 32))
            s = socket.create_connection(("127.0.0.1", port), timeout=60)
            for j in range(3):
                k = (cid * 7919 + j * 104729) % (3 * 8192)
                q, sk = _query(pir, rng, k)
                s.sendall(frame(b"\x02" + q))
                results[(cid, j)] = (k, q, sk, read_frame(s))
            s.close()
        except Exception as e:  # reported below
            errs.append(repr(e))

    ts = [threading.Thread(target=client, args=(c,)) for c in range(6)]
    for t in ts:
        t.start()
    for t in ts:
        t.join(120)
    assert not errs, errs
    assert len(results) == 18
    for (cid, j), (k, q, sk, rep) in results.items():
        assert np.array_equal(orc.unpack_coeffs_into_bytes(pir.process_reply(rep, sk))[0], _expect(buckets, k))
        if j == 0:
            oreps, _, _ = orc.process_queries(ofleet, None, q, None)
            assert rep == oreps[0], (cid, j)
    st = srv.stats()
    _say(f"[frontend] 18 queries in {st['batches'] - before} batches; front end {fe.stats()}")
    assert st["failed"] == 0


def test_a_refused_query_ends_only_its_connection(setup):
    pir, buckets, fleet, ofleet, srv, fe, port = setup
    from dnspir_amd import DnspirError
    a, b = socket.socketpair()
    rc = {}
    t = threading.Thread(target=lambda: rc.setdefault("rc", fe.handle_fd(b.fileno())))
    t.start()
    a.sendall(frame(b"\x02" + b"\x00" * 100))  # far shorter than a query
    t.join(60)
    assert rc["rc"] == -1  # DNSPIR_E_SHORT_QUERY ("not enough data", pir_wrapper.rs:124)
    a.close()
    b.close()
    # the listener still answers
    s = socket.create_connection(("127.0.0.1", port), timeout=60)
    s.sendall(frame(b"\x01"))
    assert struct.unpack(">I", read_frame(s)[1:5])[0] == ENTRIES
    s.close()
    # the advertised count must give the fleet's shape
    with pytest.raises(DnspirError) as ei:
        pir.Frontend(server=srv, db=0, num_entries=30000)  # 4 shards, the fleet has 3
    assert ei.value.code == -4
    with pytest.raises(DnspirError) as ei:
        pir.Frontend(server=srv, db=0, num_entries=300000)  # full-size shards
    assert ei.value.code == -4
