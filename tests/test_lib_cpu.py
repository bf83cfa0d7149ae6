"""CPU-side checks of the C ABI library: it loads, exports every symbol that
include/dnspir.h declares, its host-only functions (shape rule, wire sizes,
client) agree with the oracle, and device calls fail cleanly without a GPU."""
import os
import re

import numpy as np
import pytest

import pyoracle as orc
from dnspir_amd import _lib
from dnspir_amd import pir


def _declared_functions():
    text = open(_lib.HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(dnspir_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    names = _declared_functions()
    assert len(names) >= 20
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


@pytest.mark.parametrize("n", [1, 8192, 16384, 32768, 262144, 262145, 1 << 20, 3 << 20, 1 << 22])
def test_shape_rule_matches_oracle(n):
    assert pir.get_database_shape(n) == orc.database_shape(n)
    assert pir.num_databases(n) == orc.num_databases(n)
    lp, ls = pir.get_database_shape(n)
    assert pir.query_bytes(lp, ls) == orc.query_bytes(lp, ls)


def test_reply_bytes_and_required_bytes():
    assert pir.reply_bytes() == 22272
    # 26 bits per residue mod p1 and 25 mod p2: 51/64 of the reference's
    # 6*N*l*32 B slab (3328- and 3200-byte blocks of 1024 residues)
    assert _lib.lib().dnspir_required_bytes(2048) == 805306368 * 51 // 64 + 196608
    assert _lib.lib().dnspir_required_bytes(1024) == 402653184 * 51 // 64 + 196608
    assert _lib.lib().dnspir_required_bytes(512) == 6 * 2048 * 512 * 32 + 196608
    assert _lib.lib().dnspir_required_bytes(1536) == 0  # not a divisor of N
    assert _lib.lib().dnspir_required_bytes(2048 * 3 // 4) == 0


@pytest.mark.parametrize("shape", [(1024, 32), (2048, 512), (2048, 2048), (1024, 8)])
def test_client_query_bytes_match_oracle(shape):
    """prepare_query of the product and of the oracle are the same restatement:
    identical bytes and key seeds for the same RNG stream."""
    lp, ls = shape
    ro = orc.Rng(bytes([7] * 32))
    rp = pir.Rng(bytes([7] * 32))
    for (pi, si) in [(0, 0), (lp * 8 - 1, ls // 8 - 1), (2742, 0)]:
        pi %= lp * 8
        qo, sko = orc.prepare_query(ro, pi, si, lp, ls)
        qp, skp = pir.prepare_query(rp, pi, si, lp, ls)
        assert sko == skp
        assert qo == qp
        assert len(qp) == pir.query_bytes(lp, ls)


def test_client_process_reply_decrypts_oracle_reply():
    """The product's process_reply decodes a reply the oracle computed."""
    n_entries = 8192
    lp, ls = orc.database_shape(n_entries)
    data = np.random.default_rng(3).integers(0, 1024, (8 * lp, 2048)).astype(np.uint16)
    db = orc.Shard(11, lp)
    db.set_db(data)
    rng = pir.Rng(bytes(32))
    q, sk = pir.prepare_query(rng, 4321, 0, lp, ls)
    reps, _, _ = orc.process_queries([db], None, q, None)
    assert np.array_equal(pir.process_reply(reps[0], sk), data[4321])
    with pytest.raises(_lib.DnspirError):
        pir.process_reply(reps[0][:-1], sk)


def test_device_calls_without_init_fail_cleanly():
    if os.environ.get("HIP_VISIBLE_DEVICES") is None:
        import torch
        if torch.cuda.is_available():
            pytest.skip("a GPU is present")
    assert _lib.lib().dnspir_init(0) == -10
    h = __import__("ctypes").c_void_p()
    assert _lib.lib().dnspir_ctx_create(None, __import__("ctypes").byref(h)) == -10


@pytest.mark.parametrize("shape", [(1024, 32), (2048, 2048), (1024, 256)])
def test_client_keys_plus_query_equal_prepare_query(shape):
    """client_keys || client_query consume the rng as prepare_query does."""
    lp, ls = shape
    r1, r2 = pir.Rng(bytes([9] * 32)), pir.Rng(bytes([9] * 32))
    for pi, si in [(5, 0), (8 * lp - 1, ls // 8 - 1)]:
        q, sk = pir.prepare_query(r1, pi, si, lp, ls)
        k, sk2 = pir.client_keys(r2)
        assert sk == sk2 and len(k) == pir.key_bytes() == 25600
        assert k + pir.client_query(r2, sk2, pi, si, lp, ls) == q


def test_batch_scratch_bytes():
    """sizing of serving admission: grows with the batch, zero for shapes
    the batched path does not take"""
    b = [pir.batch_scratch_bytes(2048, 256, n) for n in (1, 2, 64, 65, 128)]
    assert all(x > 0 for x in b) and b == sorted(b)
    # the secondary database alone: 3 lanes x l_s columns x N u16 per query
    assert b[1] - b[0] > 3 * 2048 * 2048 * 2
    # one query on 8 or more shards reserves the two-array weights (16 MiB
    # table + tiles and corrections: 32 MiB + 16 KiB per conjugate-free stream)
    w = pir.batch_scratch_bytes(1024, 8, 1) - pir.batch_scratch_bytes(1024, 7, 1)
    assert w >= (16 << 20) + (32 << 20) + (16 << 10)
    assert pir.batch_scratch_bytes(1024, 7, 2) - pir.batch_scratch_bytes(1024, 6, 2) < (1 << 20)
    assert pir.batch_scratch_bytes(512, 4, 1) == 0      # rotated shard layout
    assert pir.batch_scratch_bytes(1024, 0, 1) == 0
    assert pir.batch_scratch_bytes(1024, 257, 1) == 0


def test_server_needs_a_device():
    import ctypes
    cfg = _lib.ServerConfig(1, 1, 1 << 20, 0)
    h = ctypes.c_void_p()
    assert _lib.lib().dnspir_server_create(ctypes.byref(cfg), ctypes.byref(h)) == -10
    bad = _lib.ServerConfig(0, 1, 1 << 20, 0)
    assert _lib.lib().dnspir_server_create(ctypes.byref(bad), ctypes.byref(h)) == -5
    bad = _lib.ServerConfig(1, 1, 1 << 20, 0, -1.0)
    assert _lib.lib().dnspir_server_create(ctypes.byref(bad), ctypes.byref(h)) in (-5, -10)


def _shard_header(l=1024, seeds=0, magic=b"DNSPIRSH", version=1, n=2048, payload=None):
    import struct
    if payload is None:
        payload = _lib.lib().dnspir_required_bytes(l)
    return struct.pack("<8sIIiiQQQQQ", magic, version, n, l, seeds, 0x1234, payload, 0, 0, 0)


def test_shard_file_header(tmp_path):
    """Shard files: the header is read without a device; wrong magic, version,
    ring size or payload size, a truncated header and a missing file are
    rejected with DNSPIR_E_FORMAT / DNSPIR_E_IO."""
    good = tmp_path / "good.shard"
    good.write_bytes(_shard_header(2048, 1))
    assert pir.shard_file_info(str(good)) == (2048, 1, _lib.lib().dnspir_required_bytes(2048))
    for name, data in [("magic", _shard_header(magic=b"DNSPIRXX")), ("version", _shard_header(version=2)),
                       ("ring", _shard_header(n=4096)), ("size", _shard_header(payload=12345)),
                       ("group", _shard_header(l=1536, payload=0)), ("short", _shard_header()[:40])]:
        f = tmp_path / f"{name}.shard"
        f.write_bytes(data)
        with pytest.raises(_lib.DnspirError) as ei:
            pir.shard_file_info(str(f))
        assert ei.value.code == -14, name
    with pytest.raises(_lib.DnspirError) as ei:
        pir.shard_file_info(str(tmp_path / "missing.shard"))
    assert ei.value.code == -13


@pytest.mark.parametrize("n", [0, 2048 * 2048 + 1, 1 << 40])
def test_shape_rule_refuses_what_the_reference_asserts(n):
    """get_database_shape asserts 1 <= num_entries <= N^2 (pir_wrapper.rs:221-222):
    the ABI returns DNSPIR_E_SHAPE and leaves the outputs untouched."""
    import ctypes
    lp, ls = ctypes.c_int(-77), ctypes.c_int(-77)
    assert _lib.lib().dnspir_database_shape(n, ctypes.byref(lp), ctypes.byref(ls)) == -4
    assert (lp.value, ls.value) == (-77, -77)
    assert _lib.lib().dnspir_num_databases(n) == 0
    with pytest.raises(_lib.DnspirError) as ei:
        pir.get_database_shape(n)
    assert ei.value.code == -4
    with pytest.raises(_lib.DnspirError) as ei:
        pir.num_databases(n)
    assert ei.value.code == -4


def test_shape_rule_edges():
    """the largest accepted database (N^2 entries = 256 full shards) and the
    secondary group's bound of N/8 shards (pir_wrapper.rs:168); zero shards
    give the smallest group, as abs_log2_ceil(0).unwrap_or(0) does there"""
    assert pir.get_database_shape(2048 * 2048) == (2048, 2048) == orc.database_shape(2048 * 2048)
    assert pir.num_databases(2048 * 2048) == 256
    assert pir.choose_index_group_secondary(256) == 2048
    assert pir.choose_index_group_secondary(0) == 8 == orc.choose_index_group_secondary(0)
    assert pir.choose_index_group_secondary(129) == 2048
    assert pir.choose_index_group_secondary(128) == 1024
    assert _lib.lib().dnspir_choose_index_group_secondary(257) == -4
    with pytest.raises(_lib.DnspirError) as ei:
        pir.choose_index_group_secondary(1 << 33)
    assert ei.value.code == -4


def test_every_error_code_has_a_text():
    """dnspir_strerror and the Python table name every DNSPIR_E_* of the header"""
    text = open(_lib.HEADER).read()
    codes = {int(v): k for k, v in re.findall(r"#define (DNSPIR_E_[A-Z_]+) \((-\d+)\)", text)}
    assert len(codes) >= 14
    for code, name in codes.items():
        s = _lib.lib().dnspir_strerror(code).decode()
        assert s and s != "unknown error", (code, name)
        assert _lib.ERRORS[code] == name
    assert _lib.lib().dnspir_strerror(0).decode() == "ok"
    assert _lib.lib().dnspir_strerror(-999).decode() == "unknown error"
    assert set(_lib.ERRORS) == set(codes)


def test_chain_cus_setter_needs_a_context():
    """dnspir_ctx_set_chain_cus without a context is DNSPIR_E_ARG (no device
    needed to refuse); the values it refuses on a context are tested on the GPU."""
    L = _lib.lib()
    for cus in (0, 1, 2, 4, 3):
        assert L.dnspir_ctx_set_chain_cus(None, cus) == -5
