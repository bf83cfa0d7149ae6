"""Pins of the oracle's primitives: ChaCha core, SEEDS, NTT convention,
mod switch, codec, shape rule and wire sizes (CPU only)."""
import hashlib

import numpy as np
import pytest

import pyoracle as orc


def _words(hexstr):
    b = bytes.fromhex(hexstr.replace(" ", ""))
    return np.frombuffer(b, dtype="<u4")


def test_chacha20_rfc8439_block():
    # RFC 8439 section 2.3.2: key 00..1f, counter 1, nonce 00:00:00:09:00:00:00:4a:00:00:00:00
    key = bytes(range(32))
    st = np.zeros(16, np.uint32)
    st[:4] = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    st[4:12] = np.frombuffer(key, dtype="<u4")
    st[12] = 1
    st[13:16] = np.frombuffer(bytes.fromhex("000000090000004a00000000"), dtype="<u4")
    out = orc.chacha_block(st, 20)
    expect = _words(
        "10f1e7e4d13b5915500fdd1fa32071c4c7d1f4c733c068030422aa9ac3d46c4e"
        "d2826446079faa0914c2d705d98b02a2b5129cd1de164eb9cbd083e8a2503c4e")
    assert np.array_equal(out, expect)


def test_chacha20_zero_key_keystream():
    # RFC 8439 appendix A.1 test vector #1 (all-zero key, nonce and counter)
    st = np.zeros(16, np.uint32)
    st[:4] = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    out = orc.chacha_block(st, 20)
    expect = _words(
        "76b8e0ada0f13d90405d6ae55386bd28bdd219b8a08ded1aa836efcc8b770dc7"
        "da41597c5157488d7724e03fb8d84a376a43b8f41518a11cc387b669b2ee6586")
    assert np.array_equal(out, expect)


def test_rng_word_order_matches_block_stream():
    """next_u64 = two consecutive keystream words, low word first; blocks in
    counter order (rand_chacha BlockRng over 4-block buffers)."""
    seed = bytes(range(100, 132))
    r = orc.Rng(seed)
    st = np.zeros(16, np.uint32)
    st[:4] = [0x61707865, 0x3320646E, 0x79622D32, 0x6B206574]
    st[4:12] = np.frombuffer(seed, dtype="<u4")
    words = []
    for ctr in range(6):
        st[12] = ctr
        words.extend(int(w) for w in orc.chacha_block(st, 12))
    got = [r.next_u32()]
    for _ in range(40):
        got.append(r.next_u64())
    assert got[0] == words[0]
    for k in range(40):
        assert got[k + 1] == words[1 + 2 * k] | (words[2 + 2 * k] << 32)


TITLES = [
    "Star Wars: Episode I - The Phantom Menace",
    "Star Wars: Episode II - Attack of the Clones",
    "Star Wars: Episode III – Revenge of the Sith",
    "Star Wars: Episode IV - A New Hope",
    "Star Wars: Episode V - The Empire Strikes Back",
    "Star Wars: Episode VI - Return of the Jedi",
]


def test_seeds_are_sha3_of_titles():
    # base_pir.rs:53-71
    for i, t in enumerate(TITLES):
        assert orc.seed(i) == hashlib.sha3_256(t.encode("utf-8")).digest()


def test_moduli_closed_form():
    assert orc.P1 == (1 << 26) - (1 << 12) + 1
    assert orc.P2 == (1 << 25) - (1 << 12) + 1
    for p in (orc.P1, orc.P2):
        assert pow(3, p - 1, p) == 1 and (p - 1) % 4096 == 0
        assert all(p % d for d in range(2, int(p ** 0.5) + 1, 1) if d < 2000)  # cheap sanity
    assert orc.Q == 2251387614257153
    # lazy window of base_pir.rs:1001-1005
    w = min((2 ** 64 - 1) // (2 * orc.P1) // (2 * orc.P1) // 3, (2 ** 64 - 1) // (2 * orc.P2) // (2 * orc.P2) // 3)
    assert w == 341


@pytest.mark.parametrize("logN", [3, 4, 6, 8])
@pytest.mark.parametrize("modulus", [0, 1])
@pytest.mark.parametrize("root_pow", [1, 3])
def test_weighted_ntt_bruteforce(logN, modulus, root_pow):
    """slot i = f(zeta^-(2i+1)) (base_pir.rs:453-467, pinned by base_pir.rs:1298)."""
    N = 1 << logN
    p = [orc.P1, orc.P2][modulus]
    rng = np.random.default_rng(logN * 10 + modulus)
    x = rng.integers(0, p, N, dtype=np.uint64)
    y, z = orc.ntt(logN, modulus, x, root_pow=root_pow)
    assert pow(z, N, p) == p - 1  # primitive 2N-th root
    zi = pow(z, p - 2, p)
    for i in range(N):
        e = (2 * i + 1)
        want = sum(int(x[j]) * pow(zi, e * j, p) for j in range(N)) % p
        assert int(y[i]) == want
    back, _ = orc.ntt(logN, modulus, y, inverse=True, root_pow=root_pow)
    assert np.array_equal(back, x)


def test_mod_switch_closed_form():
    rng = np.random.default_rng(5)
    logN = 6
    x = rng.integers(0, orc.Q, 1 << logN, dtype=np.uint64)
    x[:4] = [0, 1, orc.Q - 1, orc.Q // 2]
    for bits in (12, 17, 34, 50):
        enc = orc.mod_switch_encode(logN, bits, x)
        for xi, ei in zip(x, enc):
            c = int(xi) if int(xi) <= orc.Q // 2 else int(xi) - orc.Q
            num = c * (1 << bits)
            # round to nearest (q odd: no ties)
            r = (2 * num + orc.Q) // (2 * orc.Q)
            assert int(ei) == (r + (1 << bits)) % (1 << bits)
        dec = orc.mod_switch_decode(logN, bits, enc)
        for ei, di in zip(enc, dec):
            assert int(di) == (int(ei) * orc.Q >> bits) % orc.Q
        # decode(encode(x)) is within q / 2^(bits+1) + 1 of x
        for xi, di in zip(x, dec):
            d = (int(di) - int(xi)) % orc.Q
            d = min(d, orc.Q - d)
            assert d <= orc.Q // (1 << (bits + 1)) + 1


@pytest.mark.parametrize("bits", list(range(10, 17)) + [12, 17, 34, 50])
def test_condense_roundtrip(bits):
    # pir_wrapper.rs:442-452 uses 0..1000 for 10..16 bits; counts here are multiples of 8
    data = np.arange(1000, dtype=np.uint64) % (1 << bits)
    data = data[:1000 - 1000 % 8]
    rng = np.random.default_rng(bits)
    data = np.concatenate([data, rng.integers(0, 1 << bits, 1024, dtype=np.uint64)])
    b = orc.condense(bits, data)
    assert len(b) == data.size // 8 * bits
    back, used = orc.uncondense(bits, data.size, b)
    assert used == len(b)
    assert np.array_equal(back, data)
    with pytest.raises(ValueError):
        orc.uncondense(bits, data.size, b[:-1])


def test_condense_plane_layout():
    """12 bits: one byte plane then a nibble plane, first value in the high nibble."""
    data = np.array([0xABC, 0x123, 0x456, 0x789, 0, 0, 0, 0xFFF], np.uint64)
    b = orc.condense(12, data)
    assert b[:8] == bytes([0xBC, 0x23, 0x56, 0x89, 0, 0, 0, 0xFF])
    assert b[8:] == bytes([0xA1, 0x47, 0x00, 0x0F])
    data17 = np.array([1 << 16, 0, 0, 0, 0, 0, 0, 1 << 16 | 5], np.uint64)
    b = orc.condense(17, data17)
    assert b[16:] == bytes([0b10000001])


def test_shape_rule_and_wire_sizes():
    # pir_wrapper.rs:188,219-241 and ARTIFACT-APPENDIX.md:163-165
    assert orc.database_shape(32768) == (1024, 32)
    assert orc.num_databases(32768) == 4
    assert orc.database_shape(262144) == (1024, 256)
    assert orc.num_databases(262144) == 32
    assert orc.database_shape(262145) == (2048, 256)
    assert orc.database_shape(1 << 20) == (2048, 512)
    assert orc.num_databases(1 << 20) == 64
    assert orc.database_shape(1 << 22) == (2048, 2048)
    assert orc.num_databases(1 << 22) == 256
    assert orc.database_shape(1) == (1024, 8)
    gk = 2 * 2048 // 8 * 50
    assert gk == 25600
    assert orc.query_bytes(2048, 512) - gk == 26112
    assert orc.query_bytes(2048, 512) == 51712
    assert orc.query_bytes(2048, 2048) == 60416
    assert orc.query_bytes(1024, 32) == 43008
    assert orc.REPLY_BYTES == 22272
