"""The CPU baseline's inner product (oracle/cpu_baseline.c: main_pir_loop in
the reference's own form -- 8-lane slots side by side, lazy u64 sums reduced
every 341 rows, AVX-512 / AVX2 / scalar) gives exactly the oracle's
perform_pir_split output (base_pir.rs:958-1078) for every row loop, window,
thread count and shard shape, so the timed CPU figure is an exact one."""
import numpy as np
import pytest

import pyoracle as orc

Q = orc.Q


def _queries(N, seed, conj):
    r = np.random.default_rng(seed)
    qb = r.integers(0, Q, (8, N), dtype=np.uint64)
    qbc = r.integers(0, Q, (8, N), dtype=np.uint64) if conj else None
    gkb = r.integers(0, Q, (8, 2, N), dtype=np.uint64)
    return qb, qbc, gkb


_shards = {}


def _shard(logN, l, seed):
    key = (logN, l, seed)
    if key not in _shards:
        N = 1 << logN
        s = orc.Shard(logN, l)
        s.set_db(np.random.default_rng(seed).integers(0, 1025, (8 * l, N)).astype(np.uint16))
        _shards[key] = s
    return _shards[key]


def test_reference_window_is_341():
    """base_pir.rs:1001-1005: min over both moduli of u64::MAX / (2p)^2 / 3"""
    assert orc.cpu_reduce_every() == 341
    assert (2 ** 64 - 1) // (2 * orc.P1) ** 2 // 3 == 341
    assert orc.cpu_isa() in ("avx512", "avx2", "scalar")


@pytest.mark.parametrize("isa", [-1, 256, 0])
@pytest.mark.parametrize("l", [256, 128, 32])
@pytest.mark.parametrize("window", [0, 5])
def test_matches_perform_pir_split(isa, l, window):
    """N = 256: the full group (with its conjugate stream), half and a
    subgroup; a window of 5 rows exercises the lazy reductions these short
    loops would not reach at 341"""
    s = _shard(8, l, 60 + l)
    qb, qbc, gkb = _queries(256, l + window, s.conj)
    want = s.perform_pir_split(qb, qbc, gkb)
    orc.cpu_force_isa(isa)
    try:
        k = orc.RefK1(s, qb, qbc, gkb)
        sums = orc.RefK1.run([k], threads=3, window=window)
    finally:
        orc.cpu_force_isa(-1)
    assert np.array_equal(k.finish(sums[0]), want)


def test_fleet_of_distinct_shards_on_threads():
    """several shards in one call (their loops spread over threads), each
    shard's sums its own"""
    ss = [_shard(8, 128, 70 + k) for k in range(3)]
    qb, _, gkb = _queries(256, 5, False)
    ks = [orc.RefK1(s, qb, None, gkb) for s in ss]
    sums = orc.RefK1.run(ks, threads=4)
    for k, s, sm in zip(ks, ss, sums):
        assert np.array_equal(k.finish(sm), s.perform_pir_split(qb, None, gkb))


def test_full_size_half_shard_with_the_reference_window():
    """N = 2048, l = 1024: 1024 rows, reduced after rows 341, 682, 1023 and
    1024 as the reference does"""
    s = _shard(11, 1024, 5)
    qb, _, gkb = _queries(2048, 9, False)
    k = orc.RefK1(s, qb, None, gkb)
    sums = orc.RefK1.run([k], threads=8)
    assert np.array_equal(k.finish(sums[0]), s.perform_pir_split(qb, None, gkb))
