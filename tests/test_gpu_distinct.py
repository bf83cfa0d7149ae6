"""Byte parity of the single-query inner product on fleets of distinct shards.

Fleets of 8 or more shards take the two-array inner product (`k1w_kernel`:
the TE and IPGK1 arrays streamed against per-query weights). One query walks
the shards in groups of 16 per block, two queries (both slots of a call on
one fleet) in groups of 8. These tests run it on fleets whose shards all hold
different entries -- 20 half shards (a partial last group of 4), 17 full
shards (a last group of 1), and those 17 cycled to 129 fleet entries (the
full secondary group with its conjugate, as at 2^22 entries) -- and compare
every reply byte for byte with the oracle's (pir_wrapper.rs:349-409 over
base_pir.rs:1085-1247 and double_pir.rs:157-177), then decrypt it.
"""
import sys

import numpy as np
import pytest

import pyoracle as orc

pytestmark = pytest.mark.gpu
N = 2048


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def pir():
    from dnspir_amd import pir as p
    p.init(0)
    # a fleet that cycles shards: the oracle computes each distinct shard's
    # primary result once per query (no effect on fleets of distinct shards)
    orc.set_reuse_repeated(True)
    yield p
    orc.set_reuse_repeated(False)


_fleets = {}


def _fleet(pir, l, count):
    """`count` distinct shards of order l: GPU shards, oracle shards, entries."""
    key = (l, count)
    if key not in _fleets:
        gs, os_, es = [], [], []
        for k in range(count):
            _say(f"[distinct] preprocessing shard l={l} #{k} (GPU and oracle)")
            e = np.random.default_rng(50_000 + 97 * l + k).integers(0, 1025, (8 * l, N)).astype(np.uint16)
            g = pir.PIRDatabase(l)
            g.set_db(e)
            o = orc.Shard(11, l)
            o.set_db(e)
            gs.append(g)
            os_.append(o)
            es.append(e)
        _fleets[key] = (gs, os_, es)
    return _fleets[key]


def _two_array_bytes(pir, l, n):
    """bytes the two-array inner product streams for n shards (2/3 of the slab)"""
    return n * (pir.required_bytes(l) - 8 * N * 12) * 2 // 3


def _cases(l, n_fleet, ls):
    # (primary index, fleet position): first and last rows and lanes, the
    # last (partial) shard group, and a position in between
    return [(0, 0), (8 * l - 1, n_fleet - 1), (5 * l + 3, n_fleet // 2), (1234, min(16, n_fleet - 1)),
            (8 * l - 8, 7), (17, n_fleet - 2)]


@pytest.mark.parametrize("l,count,n_fleet", [(1024, 20, 20), (2048, 17, 17), (2048, 17, 129)])
def test_single_query_on_distinct_shards(pir, l, count, n_fleet):
    gs, os_, es = _fleet(pir, l, count)
    fg = [gs[k % count] for k in range(n_fleet)]
    fo = [os_[k % count] for k in range(n_fleet)]
    lp, ls = l, pir.choose_index_group_secondary(n_fleet)
    rng = pir.Rng(bytes([l % 251, count, n_fleet % 256] + [3] * 29))
    for t, (pi, si) in enumerate(_cases(l, n_fleet, ls)):
        q, sk = pir.prepare_query(rng, pi, si, lp, ls)
        _say(f"[distinct] l={l} fleet={n_fleet} query {t}: ({pi}, {si})")
        reps, tm = pir.process_queries(fg, None, q, None)
        assert tm.bytes_streamed == _two_array_bytes(pir, l, n_fleet), "not the two-array inner product"
        oreps, _, _ = orc.process_queries(fo, None, q, None)
        assert reps[0] == oreps[0], (l, n_fleet, t)
        assert np.array_equal(pir.process_reply(reps[0], sk), es[si % count][pi]), (l, n_fleet, t)


@pytest.mark.parametrize("l,count,n_fleet", [(1024, 20, 20), (2048, 17, 17), (2048, 17, 129)])
def test_joint_two_slot_call_on_distinct_shards(pir, l, count, n_fleet):
    """both slots on one fleet share one two-query pass (groups of 8 shards)"""
    gs, os_, es = _fleet(pir, l, count)
    fg = [gs[k % count] for k in range(n_fleet)]
    fo = [os_[k % count] for k in range(n_fleet)]
    lp, ls = l, pir.choose_index_group_secondary(n_fleet)
    rng = pir.Rng(bytes([7, l % 251, count, n_fleet % 256] + [1] * 28))
    cases = _cases(l, n_fleet, ls)
    for t in range(0, len(cases), 2):
        (p0, s0), (p1, s1) = cases[t], cases[t + 1]
        q0, sk0 = pir.prepare_query(rng, p0, s0, lp, ls)
        q1, sk1 = pir.prepare_query(rng, p1, s1, lp, ls)
        _say(f"[distinct] l={l} fleet={n_fleet} joint call {t // 2}: ({p0}, {s0}) + ({p1}, {s1})")
        reps, tm = pir.process_queries(fg, fg, q0, q1)
        assert tm.k1_launches == 1 and tm.bytes_streamed == _two_array_bytes(pir, l, n_fleet)
        oreps, _, _ = orc.process_queries(fo, fo, q0, q1)
        assert reps == oreps, (l, n_fleet, t)
        assert np.array_equal(pir.process_reply(reps[0], sk0), es[s0 % count][p0])
        assert np.array_equal(pir.process_reply(reps[1], sk1), es[s1 % count][p1])
