"""GPU parity: the HIP engine against the CPU oracle on identical seeded
inputs, bit-exact (every value is an integer mod q). Runs through the C ABI
(libdnspir.so) only."""
import sys

import numpy as np
import pytest

import pyoracle as orc

pytestmark = pytest.mark.gpu

Q = orc.Q
N = 2048


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pir():
    from dnspir_amd import pir as p
    p.init(0)
    return p


def _entries(n, seed, hi=1025):
    return np.random.default_rng(seed).integers(0, hi, (n, N)).astype(np.uint16)


def _pair(pir, l, seed, used_seeds=0, entries=None):
    e = _entries(8 * l, seed) if entries is None else entries
    g = pir.PIRDatabase(l, used_seeds)
    g.set_db(e)
    o = orc.Shard(11, l, used_seeds=used_seeds)
    o.set_db(e)
    return g, o, e


_cache = {}


def _shard(pir, l, seed=1, used_seeds=0):
    key = (l, seed, used_seeds)
    if key not in _cache:
        _say(f"[parity] preprocessing l={l} seeds={used_seeds} seed={seed}")
        _cache[key] = _pair(pir, l, seed, used_seeds)
    return _cache[key]


def _rand_q(seed, n=1):
    r = np.random.default_rng(seed)
    return r.integers(0, Q, (n, N), dtype=np.uint64)


@pytest.mark.parametrize("l,used_seeds", [(1024, 0), (2048, 0), (8, 1), (64, 1), (256, 0), (2048, 1)])
def test_get_a_and_perform_pir(pir, l, used_seeds):
    g, o, _ = _shard(pir, l, 1, used_seeds)
    _say(f"[parity] get_a l={l}")
    assert np.array_equal(g.get_a(), o.get_a())
    conj = l == N
    for seed in (10, 11):
        qb = _rand_q(seed)[0]
        qbc = _rand_q(seed + 100)[0] if conj else None
        gk = _rand_q(seed + 200, 2)
        _say(f"[parity] perform_pir l={l} seed={seed}")
        assert np.array_equal(g.perform_pir(qb, qbc, gk), o.perform_pir(qb, qbc, gk))


@pytest.mark.parametrize("l", [1024, 2048, 32])
def test_perform_pir_split(pir, l):
    g, o, _ = _shard(pir, l, 1, 0)
    qb = _rand_q(30, 8)
    qbc = _rand_q(31, 8) if l == N else None
    gk = _rand_q(32, 16).reshape(8, 2, N)
    _say(f"[parity] perform_pir_split l={l}")
    assert np.array_equal(g.perform_pir_split(qb, qbc, gk), o.perform_pir_split(qb, qbc, gk))


def test_edge_entries_all_zero_and_all_max(pir):
    for val in (0, 1024):
        e = np.full((8 * 1024, N), val, np.uint16)
        g, o, _ = _pair(pir, 1024, 0, entries=e)
        qb = _rand_q(40)[0]
        gk = _rand_q(41, 2)
        _say(f"[parity] constant entries {val}")
        assert np.array_equal(g.get_a(), o.get_a())
        assert np.array_equal(g.perform_pir(qb, None, gk), o.perform_pir(qb, None, gk))


def _block(seed, conj_p, conj_s):
    q = _rand_q(seed, 6)
    if not conj_p:
        q[3] = 0
    if not conj_s:
        q[5] = 0
    return q


@pytest.mark.parametrize("l,lens", [(1024, (1, 0)), (1024, (3, 2)), (2048, (2, 2)), (64, (5, 1))])
def test_double_pir(pir, l, lens):
    g, o, _ = _shard(pir, l, 1, 0)
    g2, o2, _ = _shard(pir, l, 2, 0)
    fg = [g, g2, g][:lens[0]] if lens[0] <= 3 else [g, g2] * (lens[0] // 2) + [g] * (lens[0] % 2)
    fo = [o, o2, o][:lens[0]] if lens[0] <= 3 else [o, o2] * (lens[0] // 2) + [o] * (lens[0] % 2)
    ls = orc.choose_index_group_secondary(lens[0])
    conj_p, conj_s = l == N, ls == N
    q0 = _block(50, conj_p, conj_s)
    q1 = _block(51, conj_p, conj_s) if lens[1] else None
    _say(f"[parity] double_pir l={l} lens={lens}")
    go, tm = pir.perform_batched_double_pir(fg, fg[:lens[1]] if lens[1] else None, q0, q1)
    oo, _, _ = orc.double_pir(fo, fo[:lens[1]] if lens[1] else None, q0, q1, ls)
    assert np.array_equal(go[0], oo[0])
    if lens[1]:
        assert np.array_equal(go[1], oo[1])
        # the same query alone gives the same bytes as in the mix
        alone, _ = pir.perform_batched_double_pir(fg, None, q0, None)
        assert np.array_equal(alone[0], go[0])


def test_double_pir_full_group_secondary(pir):
    """More than N/16 shards: full-group secondary with conjugated query
    (the 2^22 shape), fleet of 130 repeated half shards."""
    g, o, _ = _shard(pir, 1024, 1, 0)
    g2, o2, _ = _shard(pir, 1024, 2, 0)
    fg = [g, g2] * 65
    fo = [o, o2] * 65
    ls = orc.choose_index_group_secondary(len(fo))
    assert ls == N
    q0 = _block(60, False, True)
    _say("[parity] double_pir full-group secondary, 130 shards")
    go, _ = pir.perform_batched_double_pir(fg, None, q0, None)
    oo, _, _ = orc.double_pir(fo, None, q0, None, ls)
    assert np.array_equal(go[0], oo[0])


@pytest.mark.parametrize("l", [1024, 2048])
def test_process_queries_bytes_and_decrypt(pir, l):
    g, o, e = _shard(pir, l, 1, 0)
    g2, o2, e2 = _shard(pir, l, 2, 0)
    fg, fo = [g, g2], [o, o2]
    lp = l
    ls = orc.choose_index_group_secondary(2)
    rng = pir.Rng(bytes([5] * 32))
    for (pi, si) in [(0, 0), (8 * l - 1, 1), (2742 % (8 * l), 1)]:
        q, sk = pir.prepare_query(rng, pi, si, lp, ls)
        _say(f"[parity] process_queries l={l} idx=({pi},{si})")
        reps, tm = pir.process_queries(fg, None, q, None)
        oreps, _, _ = orc.process_queries(fo, None, q, None)
        assert reps[0] == oreps[0]
        want = (e, e2)[si][pi]
        assert np.array_equal(pir.process_reply(reps[0], sk), want)


@pytest.mark.parametrize("l", [1024, 2048])
def test_two_stream_inner_product_matches_three_streams(pir, l):
    """The single-query inner product over the TE and IPGK1 arrays only (IPGK0
    folded into per-query weights through the key-switch identity) gives the
    oracle's bytes, like the three-array one, on both slots (one two-query
    pass, as both slots use the same fleet); the automatic choice agrees."""
    g, o, e = _shard(pir, l, 1, 0)
    g2, o2, e2 = _shard(pir, l, 2, 0)
    fg, fo = [g, g2, g], [o, o2, o]
    ls = orc.choose_index_group_secondary(3)
    rng = pir.Rng(bytes([7] * 32))
    q0, sk0 = pir.prepare_query(rng, 8 * l - 1, 1, l, ls)
    q1, sk1 = pir.prepare_query(rng, 5, 2, l, ls)
    oreps, _, _ = orc.process_queries(fo, fo, q0, q1)
    got = {}
    for mode in (2, 3, 0):
        ctx = pir.Context()
        ctx.set_k1_mode(mode)
        _say(f"[parity] two-stream K1 l={l} mode={mode}")
        got[mode], tm = pir.process_queries(fg, fg, q0, q1, ctx=ctx)
        if mode == 2:  # one pass for both slots
            assert tm.bytes_streamed * 3 == 2 * 3 * _lib_slab_bytes(pir, l)
            assert tm.k1_launches == 1
    assert got[2] == oreps and got[3] == oreps and got[0] == oreps
    assert np.array_equal(pir.process_reply(got[2][0], sk0), e2[8 * l - 1])
    assert np.array_equal(pir.process_reply(got[2][1], sk1), e[5])
    # raw ring elements of the composition, no wire rounding in between
    ctx = pir.Context()
    ctx.set_k1_mode(2)
    qb = _block(70, l == N, ls == N)
    go, _ = pir.perform_batched_double_pir(fg, None, qb, None, ctx=ctx)
    oo, _, _ = orc.double_pir(fo, None, qb, None, ls)
    assert np.array_equal(go[0], oo[0])


def _lib_slab_bytes(pir, l):
    from dnspir_amd import lib
    return lib().dnspir_required_bytes(l) - 196608


def test_process_queries_two_slots(pir):
    g, o, e = _shard(pir, 1024, 1, 0)
    rng = pir.Rng(bytes([6] * 32))
    q0, sk0 = pir.prepare_query(rng, 77, 0, 1024, 8)
    q1, sk1 = pir.prepare_query(rng, 4000, 0, 1024, 8)
    _say("[parity] process_queries two slots")
    reps, _ = pir.process_queries([g], [g], q0, q1)
    oreps, _, _ = orc.process_queries([o], [o], q0, q1)
    assert reps == oreps
    assert np.array_equal(pir.process_reply(reps[0], sk0), e[77])
    assert np.array_equal(pir.process_reply(reps[1], sk1), e[4000])


def test_error_paths(pir):
    from dnspir_amd import DnspirError
    g, _, e = _shard(pir, 1024, 1, 0)
    bad = pir.PIRDatabase(1024)
    with pytest.raises(DnspirError) as ei:
        bad.set_db(e[:-1])
    assert ei.value.code == -3
    full, _, _ = _shard(pir, 2048, 1, 0)
    with pytest.raises(DnspirError) as ei:
        full.perform_pir(_rand_q(1)[0], None, _rand_q(2, 2))
    assert ei.value.code == -2
    rng = pir.Rng(bytes(32))
    q, _ = pir.prepare_query(rng, 1, 0, 1024, 8)
    with pytest.raises(DnspirError) as ei:
        pir.process_queries([g], None, q[:-1], None)
    assert ei.value.code == -1


@pytest.mark.parametrize("cus", [0, 2])
def test_chain_exchange_timeout_runs_again_on_one_cu(pir, cus):
    """The multi-CU secondary chain (four CUs per chain by default for a few
    chains, two when the context caps it) bounds every wait for its partner
    blocks: with the bound forced to zero from the host every wait gives up at
    once (every kernel still completes), the call runs its secondary phase
    again with one CU per chain and still gives the oracle's bytes; the
    context counts the retry. With the bound restored the same context
    answers on several CUs."""
    g, o, e = _shard(pir, 1024, 1, 0)
    rng = pir.Rng(bytes([9] * 32))
    q, sk = pir.prepare_query(rng, 123, 0, 1024, 8)
    q2, sk2 = pir.prepare_query(rng, 4095, 0, 1024, 8)
    oreps, _, _ = orc.process_queries([o], None, q, None)
    ctx = pir.Context()
    ctx.set_chain_cus(cus)
    ctx.set_exchange_polls(0)
    _say(f"[parity] chain exchange bound 0, chain CUs {cus}: expect a retry on one CU")
    reps, _ = pir.process_queries([g], None, q, None, ctx=ctx)
    assert reps[0] == oreps[0]
    st = ctx.stats()
    assert st["pair_retries"] == 1 and st["pair_runs"] == 1, st
    assert st["quad_runs"] == (1 if cus == 0 else 0), st
    # a two-query batch (graph captured with the two-CU chain) takes the same retry
    breps, _ = pir.process_batch([g], [q, q2], ctx=ctx)
    assert breps[0] == oreps[0]
    assert np.array_equal(pir.process_reply(breps[1], sk2), e[4095])
    assert ctx.stats()["pair_retries"] == 2
    ctx.set_exchange_polls(1 << 21)
    reps, _ = pir.process_queries([g], None, q, None, ctx=ctx)
    assert reps[0] == oreps[0]
    assert np.array_equal(pir.process_reply(reps[0], sk), e[123])
    st = ctx.stats()
    assert st["pair_retries"] == 2 and st["pair_runs"] == 3, st
    assert st["quad_runs"] == (3 if cus == 0 else 0), st


@pytest.mark.parametrize("nsh,two", [(1, False), (3, True), (17, False), (130, False), (130, True)])
def test_secondary_chain_on_one_two_four_cus(pir, nsh, two):
    """The secondary chain on one CU per chain, on two (one per modulus) and
    on four (one per modulus and half of the slots, which sigma_5 keeps
    apart) gives the oracle's ring elements at secondary lengths 8, 32, 256
    and the conjugated 2048 (3 to 12 chains), alone and with a second slot."""
    g, o, _ = _shard(pir, 1024, 1, 0)
    g2, o2, _ = _shard(pir, 1024, 2, 0)
    fg = ([g, g2] * nsh)[:nsh]
    fo = ([o, o2] * nsh)[:nsh]
    ls = orc.choose_index_group_secondary(nsh)
    q0 = _block(80 + nsh, False, ls == N)
    q1 = _block(90 + nsh, False, ls == N) if two else None
    f1g, f1o = (fg[:2], fo[:2]) if two else (None, None)
    oo, _, _ = orc.double_pir(fo, f1o, q0, q1, ls)
    for cus in (1, 2, 4):
        ctx = pir.Context()
        ctx.set_chain_cus(cus)
        _say(f"[parity] secondary chain on {cus} CU(s), {nsh} shards (l_s={ls}), two slots {two}")
        go, _ = pir.perform_batched_double_pir(fg, f1g, q0, q1, ctx=ctx)
        assert np.array_equal(go[0], oo[0]), cus
        if two:
            assert np.array_equal(go[1], oo[1]), cus
        st = ctx.stats()
        assert st["pair_runs"] == (0 if cus == 1 else 1) and st["quad_runs"] == (1 if cus == 4 else 0), st
        assert st["pair_retries"] == 0, st


def test_chain_cus_setting_is_checked(pir):
    from dnspir_amd import DnspirError
    ctx = pir.Context()
    for bad in (3, 5, 8, -1):
        with pytest.raises(DnspirError) as ei:
            ctx.set_chain_cus(bad)
        assert ei.value.code == -5
    for good in (0, 1, 2, 4):
        ctx.set_chain_cus(good)


def test_batch_on_four_cu_chains_matches_alone(pir):
    """A batch of five queries (15 chains: four-CU blocks for 16, the last
    chain's blocks idle) answers every query as process_queries does alone."""
    g, o, e = _shard(pir, 1024, 1, 0)
    rng = pir.Rng(bytes([15] * 32))
    idx = [0, 1, 4097, 8191, 777]
    qs = [pir.prepare_query(rng, k, 0, 1024, 8) for k in idx]
    ctx = pir.Context()
    _say("[parity] batch of 5 on four-CU chains")
    reps, _ = pir.process_batch([g], [q for q, _ in qs], ctx=ctx)
    assert ctx.stats()["quad_runs"] == 1
    for k, (q, sk), rep in zip(idx, qs, reps):
        oreps, _, _ = orc.process_queries([o], None, q, None)
        assert rep == oreps[0], k
        assert np.array_equal(pir.process_reply(rep, sk), e[k])


def test_fleet_pointer_list_follows_new_shards(pir):
    """A context caches the device list of a fleet's shard pointers. Shard
    handles freed and made again can come back at the same host addresses for
    shards in other device memory: the next call must read the new shards."""
    import gc
    rng = pir.Rng(bytes([5] * 32))
    ctx = pir.Context()
    answers = []
    for rnd in range(3):
        e = [np.random.default_rng(900 + 10 * rnd + k).integers(0, 1025, (8 * 1024, N)).astype(np.uint16)
             for k in range(2)]
        keep = [pir.PIRDatabase(1024) for _ in range(rnd)]  # shifts the device memory of the next shards
        fl = [pir.PIRDatabase(1024) for _ in range(2)]
        for g, x in zip(fl, e):
            g.set_db(x)
        idx = 8 * 1024 - 1 - rnd
        q, sk = pir.prepare_query(rng, idx, 1, 1024, 16)
        reps, _ = pir.process_queries(fl, None, q, None, ctx=ctx)
        got = pir.process_reply(reps[0], sk)
        _say(f"[parity] fleet round {rnd}: handles {[g.h.value for g in fl]}")
        assert np.array_equal(got, e[1][idx]), rnd
        answers.append(got)
        del fl, keep, g
        gc.collect()
    assert not np.array_equal(answers[0], answers[1])
