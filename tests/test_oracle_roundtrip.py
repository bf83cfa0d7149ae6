"""Decryption round trips of the oracle, restating the reference's own tests
base_pir.rs:1413-1537 and double_pir.rs:328-451 (N = 256), plus
root-of-unity independence of the response and the N = 2048 byte API."""
import numpy as np
import pytest

import pyoracle as orc

T = orc.T


def _mono(N, e):
    m = np.zeros(N, np.uint16)
    e %= 2 * N
    if e < N:
        m[e] = 1
    else:
        m[e - N] = T - 1
    return m


def _const_entries(n, N):
    e = np.zeros((n, N), np.uint16)
    e[:, 0] = np.arange(n) % T
    return e


def test_base_pir_single():
    # base_pir.rs:1413-1452: full group, data i as constant polynomial
    logN, N = 8, 256
    rng = orc.Rng(bytes(32))
    db = orc.Shard(logN, N)
    data = _const_entries(8 * N, N)
    db.set_db(data)
    sk = orc.gen_sk(logN, rng)
    gk = orc.gen_gk_b(logN, rng, sk)
    a = db.get_a()
    for idx in range(32):
        qb = orc.enc_sym_b(logN, orc.seed(2), rng, sk, _mono(N, idx))
        qbc = orc.enc_sym_b(logN, orc.seed(3), rng, sk, _mono(N, 2 * N - idx))
        b = db.perform_pir(qb, qbc, gk)
        res = orc.dec(logN, sk, a[0], b[0])
        assert np.array_equal(res, data[idx * 8])


def test_base_pir_nontrivial_quotient():
    # base_pir.rs:1454-1494: subgroup of order 32, no conjugate
    logN, N, l = 8, 256, 32
    rng = orc.Rng(bytes(32))
    db = orc.Shard(logN, l)
    assert not db.conj
    data = _const_entries(8 * l, N)
    db.set_db(data)
    sk = orc.gen_sk(logN, rng)
    gk = orc.gen_gk_b(logN, rng, sk)
    a = db.get_a()
    k = N // 2 // l
    for idx in range(32):
        qb = orc.enc_sym_b(logN, orc.seed(2), rng, sk, _mono(N, idx * k))
        b = db.perform_pir(qb, None, gk)
        assert np.array_equal(orc.dec(logN, sk, a[0], b[0]), data[idx * 8])


def test_base_pir_split():
    # base_pir.rs:1496-1537: 8 keys, one per lane, random data
    logN, N = 8, 256
    rng = orc.Rng(bytes(32))
    db = orc.Shard(logN, N)
    data = np.random.default_rng(3).integers(0, T, (8 * N, N)).astype(np.uint16)
    db.set_db(data)
    sks = [orc.gen_sk(logN, rng) for _ in range(8)]
    gks = np.stack([orc.gen_gk_b(logN, rng, sks[i]) for i in range(8)])
    a = db.get_a()
    for idx in (0, 1, 17, 31):
        qb = np.stack([orc.enc_sym_b(logN, orc.seed(2), rng, sks[i], _mono(N, idx)) for i in range(8)])
        qbc = np.stack([orc.enc_sym_b(logN, orc.seed(3), rng, sks[i], _mono(N, 2 * N - idx)) for i in range(8)])
        b = db.perform_pir_split(qb, qbc, gks)
        for i in range(8):
            assert np.array_equal(orc.dec(logN, sks[i], a[i], b[i]), data[idx * 8 + i])


def _enc_double(logN, rng, sk, lp, ls, pidx, sidx):
    """enc_double_pir, double_pir.rs:245-290; returns a [6][N] query block."""
    N = 1 << logN
    kp = N // 2 // (lp if lp < N else N // 2)
    ks = N // 2 // (ls if ls < N else N // 2)
    i = pidx // 8
    j = sidx * 8 + pidx % 8
    gk = orc.gen_gk_b(logN, rng, sk)
    q = np.zeros((6, N), np.uint64)
    q[0:2] = gk
    q[2] = orc.enc_sym_b(logN, orc.seed(2), rng, sk, _mono(N, i * kp))
    if lp == N:
        q[3] = orc.enc_sym_b(logN, orc.seed(3), rng, sk, _mono(N, 2 * N - i * kp))
    q[4] = orc.enc_sym_b(logN, orc.seed(4), rng, sk, _mono(N, j * ks))
    if ls == N:
        q[5] = orc.enc_sym_b(logN, orc.seed(5), rng, sk, _mono(N, 2 * N - j * ks))
    return q


def _dec_double(logN, sk, reply):
    """dec_double_pir, double_pir.rs:292-317"""
    parts = [orc.dec(logN, sk, reply[p, 0], reply[p, 1]).astype(np.int64) for p in range(3)]
    v = parts[2] * 1025 * 1025 + parts[1] * 1025 + parts[0]
    assert np.all((v >> 12) < (1 << 17))
    a = orc.mod_switch_decode(logN, 17, (v >> 12).astype(np.uint64))
    b = orc.mod_switch_decode(logN, 12, (v & 4095).astype(np.uint64))
    return orc.dec(logN, sk, a, b)


def _seeded_entries(n, N, offset=0):
    """entry i: seeded uniform coefficients mod t (stand-in for the
    reference's per-entry seeded random_element)."""
    out = np.zeros((n, N), np.uint16)
    for i in range(n):
        out[i] = np.random.default_rng(offset + i).integers(0, T, N)
    return out


def test_double_pir():
    # double_pir.rs:328-388: one full-group shard, secondary full group,
    # including a secondary index beyond the fleet (decrypts to zero)
    logN, N = 8, 256
    rng = orc.Rng(bytes(32))
    data = _seeded_entries(8 * N, N)
    db = orc.Shard(logN, N)
    db.set_db(data)
    sk = orc.gen_sk(logN, rng)
    for (i, j) in [(0, 0), (1, 1), (2, 0), (2, 1)]:
        q = _enc_double(logN, rng, sk, N, N, i, j)
        outs, _, _ = orc.double_pir([db], None, q, None, N)
        res = _dec_double(logN, sk, outs[0])
        if j < 1:
            assert np.array_equal(res, data[i + j * N])
        else:
            assert not res.any()


def test_double_pir_primary_subgroup():
    # double_pir.rs:395-451: half-size primary shards, 2 shards, subgroup secondary
    logN, N = 8, 256
    lp, nshards = N // 2, 2
    ls = 8 * nshards
    rng = orc.Rng(bytes([1] * 32))
    per = 8 * lp
    data = _seeded_entries(nshards * per, N)
    dbs = []
    for s in range(nshards):
        d = orc.Shard(logN, lp)
        d.set_db(data[s * per:(s + 1) * per])
        dbs.append(d)
    sk = orc.gen_sk(logN, rng)
    for (i, j) in [(0, 0), (1, 1), (2, 0), (9, 1), (per - 1, 0), (per - 1, 1)]:
        q = _enc_double(logN, rng, sk, lp, ls, i, j)
        outs, _, _ = orc.double_pir(dbs, None, q, None, ls)
        assert np.array_equal(_dec_double(logN, sk, outs[0]), data[i + j * per])


def test_double_pir_two_slots_independent():
    """Both batch slots in one call give the same bytes as each alone
    (lanes never interact)."""
    logN, N = 8, 256
    lp = N // 2
    rng = orc.Rng(bytes([2] * 32))
    data = _seeded_entries(8 * lp, N, offset=500)
    db = orc.Shard(logN, lp)
    db.set_db(data)
    ls = 8
    sk0 = orc.gen_sk(logN, rng)
    sk1 = orc.gen_sk(logN, rng)
    q0 = _enc_double(logN, rng, sk0, lp, ls, 5, 0)
    q1 = _enc_double(logN, rng, sk1, lp, ls, 77, 0)
    both, _, _ = orc.double_pir([db], [db], q0, q1, ls)
    alone0, _, _ = orc.double_pir([db], None, q0, None, ls)
    alone1, _, _ = orc.double_pir([db], [db], None, q1, ls)
    assert np.array_equal(both[0], alone0[0])
    assert np.array_equal(both[1], alone1[1])
    assert np.array_equal(_dec_double(logN, sk0, both[0]), data[5])
    assert np.array_equal(_dec_double(logN, sk1, both[1]), data[77])


def test_response_independent_of_root_choice():
    """The response is a ring-level function: a different primitive 2N-th
    root (zeta^3, zeta^5) must give identical values."""
    logN, N = 8, 256
    lp = N // 2
    rng = orc.Rng(bytes([9] * 32))
    data = _seeded_entries(8 * lp, N, offset=1000)
    sk = orc.gen_sk(logN, rng)
    q = _enc_double(logN, rng, sk, lp, 8, 33, 0)
    res = []
    for rp in (1, 3, 5):
        db = orc.Shard(logN, lp, root_pow=rp)
        db.set_db(data)
        outs, _, _ = orc.double_pir([db], None, q, None, 8)
        res.append(outs[0])
    assert np.array_equal(res[0], res[1]) and np.array_equal(res[0], res[2])
    assert np.array_equal(_dec_double(logN, sk, res[0]), data[33])


@pytest.mark.slow
def test_process_queries_bytes_n2048_half_shards():
    """Byte API at N = 2048: two half-size shards (16384 entries),
    prepare_query -> process_queries -> process_reply, reply size 22272 B."""
    n_entries = 16384
    lp, ls = orc.database_shape(n_entries)
    assert (lp, ls) == (1024, 16)
    nsh = orc.num_databases(n_entries)
    per = 8 * lp
    data = np.random.default_rng(11).integers(0, 1024, (nsh * per, 2048)).astype(np.uint16)
    dbs = []
    for s in range(nsh):
        d = orc.Shard(11, lp)
        d.set_db(data[s * per:(s + 1) * per])
        dbs.append(d)
    rng = orc.Rng(bytes(32))
    for (pi, si) in [(0, 0), (8191, 1)]:
        query, sk = orc.prepare_query(rng, pi, si, lp, ls)
        assert len(query) == orc.query_bytes(lp, ls)
        reps, _, _ = orc.process_queries(dbs, None, query, None)
        assert len(reps[0]) == 22272
        assert np.array_equal(orc.process_reply(reps[0], sk), data[si * per + pi])
    with pytest.raises(ValueError):
        orc.process_queries(dbs, None, query[:-1], None)
