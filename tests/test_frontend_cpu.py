"""The DNS-over-PIR front end on the CPU (no GPU, so no PIR fleet): framing,
INFO / QUERY / DUMP answers and the connection errors of the reference's
protocol (crates/dns-over-pir/src/dns/protocol.rs, server.rs:747-807), over a
socket pair and over the TCP listener; and the bucket packing
(dns/bucket.rs:271-303) against the oracle and the reference's own round-trip
fixture (bucket.rs:619-640)."""
import socket
import struct
import threading

import numpy as np
import pytest

import pyoracle as orc
from dnspir_amd import pir

FRAME_MAX = 64 << 20


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


def decode_info(p):
    """decode_info, protocol.rs:179-200"""
    flags = p[0]
    num_entries, zl = struct.unpack(">II", p[1:9])
    zone = p[9:9 + zl].decode()
    dump_size, salt = struct.unpack(">IQ", p[9 + zl:9 + zl + 12])
    assert len(p) == 9 + zl + 12
    return flags, num_entries, zone, dump_size, salt


class Conn:
    """one connection handled by dnspir_frontend_handle_fd on a thread"""

    def __init__(self, fe):
        self.c, srv = socket.socketpair()
        self.srv = srv
        self.rc = None
        self.t = threading.Thread(target=self._run, args=(fe,))
        self.t.start()

    def _run(self, fe):
        self.rc = fe.handle_fd(self.srv.fileno())

    def end(self):
        self.c.close()
        self.t.join(10)
        assert not self.t.is_alive()
        self.srv.close()
        return self.rc


def test_info_query_dump_without_pir():
    dump = bytes(range(256)) * 3
    fe = pir.Frontend(zone_name="org", hash_salt=0xDEADBEEFCAFEF00D, dump=dump)
    k = Conn(fe)
    k.c.sendall(frame(b"\x01"))
    # the reference's info_roundtrip fixture (protocol.rs:206-222), without PIR
    assert decode_info(read_frame(k.c)) == (2, 0, "org", len(dump), 0xDEADBEEFCAFEF00D)
    k.c.sendall(frame(b"\x02" + b"\x00" * 100))
    assert read_frame(k.c) == b"\xff"  # RESP_UNAVAILABLE: no PIR here
    k.c.sendall(frame(b"\x03"))
    assert read_frame(k.c) == dump
    k.c.sendall(frame(b"\x01") + frame(b"\x03"))  # two requests back to back
    assert decode_info(read_frame(k.c))[2] == "org"
    assert read_frame(k.c) == dump
    assert k.end() == 0  # the peer closed between commands
    st = fe.stats()
    assert st["frames_in"] == 5 and st["frames_out"] == 5 and st["infos"] == 2 and st["dumps"] == 2
    assert st["unavailable"] == 1 and st["bytes_out"] == sum(4 + n for n in (24, 1, 768, 24, 768))  # INFO: 1 + 4 + 4 + 3 + 4 + 8


def test_empty_zone_and_no_dump():
    """info_empty_zone_name (protocol.rs:225-238) and DUMP without a dump"""
    fe = pir.Frontend(zone_name="", hash_salt=0)
    k = Conn(fe)
    k.c.sendall(frame(b"\x01"))
    assert decode_info(read_frame(k.c)) == (0, 0, "", 0, 0)
    k.c.sendall(frame(b"\x03"))
    assert read_frame(k.c) == b"\xff"
    assert k.end() == 0


@pytest.mark.parametrize("case", ["oversized", "empty", "unknown", "unknown_after_info"])
def test_bad_requests_end_the_connection(case):
    fe = pir.Frontend(zone_name="x")
    k = Conn(fe)
    if case == "oversized":
        k.c.sendall(struct.pack(">I", FRAME_MAX + 1))  # rejected before any payload is read
    elif case == "empty":
        k.c.sendall(frame(b""))
    elif case == "unknown":
        k.c.sendall(frame(b"\x09abc"))
    else:
        k.c.sendall(frame(b"\x01"))
        assert decode_info(read_frame(k.c))[2] == "x"
        k.c.sendall(frame(b"\x00"))
    k.t.join(10)
    assert k.rc == -16
    k.srv.close()  # handle_fd leaves the fd open; nothing more was answered on it
    k.c.settimeout(5)
    with pytest.raises((EOFError, ConnectionError)):
        read_frame(k.c)
    k.c.close()


def test_largest_frame_is_accepted_and_cut_frames_end_cleanly():
    fe = pir.Frontend(zone_name="z", dump=b"D")
    # a 64 MiB frame is allowed (protocol.rs:111 rejects only larger ones);
    # an unknown command inside it still ends the connection
    k = Conn(fe)
    big = b"\x07" + bytes(FRAME_MAX - 1)
    k.c.sendall(frame(big))
    k.t.join(30)
    assert k.rc == -16
    k.end()
    # a frame cut short ends the connection like a close (read_exact's UnexpectedEof)
    for cut in (b"\x00\x00", struct.pack(">I", 10) + b"\x01\x02\x03"):
        k = Conn(fe)
        k.c.sendall(cut)
        assert k.end() == 0


def test_listener_serves_connections_on_threads():
    fe = pir.Frontend(zone_name="net", hash_salt=7, dump=b"zone-bytes")
    port = fe.listen("127.0.0.1", 0)
    assert port > 0
    socks = []
    for i in range(3):
        s = socket.create_connection(("127.0.0.1", port), timeout=10)
        s.sendall(frame(b"\x01"))
        assert decode_info(read_frame(s)) == (2, 0, "net", 10, 7)
        socks.append(s)
    socks[1].sendall(frame(b"\x03"))
    assert read_frame(socks[1]) == b"zone-bytes"
    socks[2].sendall(frame(b"\x05"))  # unknown command: only this connection ends
    with pytest.raises((EOFError, ConnectionError)):
        read_frame(socks[2])
    socks[0].sendall(frame(b"\x01"))
    assert decode_info(read_frame(socks[0]))[4] == 7
    for s in socks:
        s.close()
    st = fe.stats()
    assert st["infos"] == 4 and st["dumps"] == 1
    fe.close()  # stops the listener and joins every connection thread


def test_listener_counts_errors():
    fe = pir.Frontend(zone_name="a")
    port = fe.listen("127.0.0.1", 0)
    s = socket.create_connection(("127.0.0.1", port), timeout=10)
    s.sendall(frame(b""))
    with pytest.raises((EOFError, ConnectionError)):
        read_frame(s)
    s.close()
    s = socket.create_connection(("127.0.0.1", port), timeout=10)
    s.sendall(frame(b"\x01"))
    read_frame(s)
    s.close()
    for _ in range(100):
        st = fe.stats()
        if st["connections"] == 2 and st["errors"] == 1:
            break
        threading.Event().wait(0.05)
    assert st["connections"] == 2 and st["errors"] == 1
    fe.close()


def test_frontend_config_checks():
    from dnspir_amd import DnspirError
    # a listener that cannot bind reports DNSPIR_E_IO
    fe = pir.Frontend()
    port = fe.listen("127.0.0.1", 0)
    fe2 = pir.Frontend()
    with pytest.raises(DnspirError) as ei:
        fe2.listen("127.0.0.1", port)
    assert ei.value.code == -13
    with pytest.raises(DnspirError) as ei:
        fe2.listen("not-an-address", 0)
    assert ei.value.code == -5
    fe.close()
    fe2.close()


def test_pack_matches_the_reference_fixture_and_the_oracle():
    """coeff_byte_roundtrip (bucket.rs:619-640): bucket[i] = (i*31 + 7) & 0xFF
    packs into coefficients below 2^10 and unpacks back; the product's host
    packing equals the oracle's; the first group is checked by hand"""
    bucket = np.array([(i * 31 + 7) & 0xFF for i in range(2560)], np.uint8)
    c = pir.pack_bytes_into_coeffs(bucket[None])
    assert c.shape == (1, 2048) and int(c.max()) < 1024
    assert np.array_equal(c, orc.pack_bytes_into_coeffs(bucket[None]))
    assert np.array_equal(orc.unpack_coeffs_into_bytes(c)[0], bucket)
    # bytes 07 26 45 64 83 read little endian: v = 0x8364452607
    # = 10 0000 1101 | 10 0100 0100 | 01 0100 1001 | 10 0000 0111 (bits 39..0)
    assert list(c[0, :4]) == [0x207, 0x149, 0x244, 0x20D] == [519, 329, 580, 525]
    rnd = np.random.default_rng(5).integers(0, 256, (7, 2560), dtype=np.uint8)
    pc = pir.pack_bytes_into_coeffs(rnd)
    assert np.array_equal(pc, orc.pack_bytes_into_coeffs(rnd))
    assert np.array_equal(orc.unpack_coeffs_into_bytes(pc), rnd)
    assert not pir.pack_bytes_into_coeffs(np.zeros((2, 2560), np.uint8)).any()
