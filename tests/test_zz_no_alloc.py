"""dnspir_process_queries_noalloc: the no-allocation variant of
dnspir_process_queries_device (all scratch in the library's fixed per-GPU
array) answers byte for byte as the existing entry point, call after call on
the same array (the second call smaller than the first), allocates no device
memory, and refuses a call whose scratch does not fit before any device work."""
import ctypes
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


def _shards(pir, l, count, seed):
    out = []
    for k in range(count):
        e = np.random.default_rng(seed + k).integers(0, 1025, (8 * l, N)).astype(np.uint16)
        g = pir.PIRDatabase(l)
        g.set_db(e)
        out.append((g, e))
    return out


def _dev(torch, b):
    return torch.tensor(np.frombuffer(bytes(b), np.uint8).copy(), device="cuda")


def _both(pir, torch, fleet0, fleet1, q0, q1):
    """replies of the existing entry point and of the no-allocation one"""
    rb = pir.reply_bytes()
    qd = [_dev(torch, q) if q is not None else None for q in (q0, q1)]
    want = [torch.zeros(rb, dtype=torch.uint8, device="cuda") if q is not None else None for q in (q0, q1)]
    got = [torch.zeros(rb, dtype=torch.uint8, device="cuda") if q is not None else None for q in (q0, q1)]
    pir.process_queries_device(fleet0, fleet1, qd[0], qd[1], want[0], want[1])
    free0 = torch.cuda.mem_get_info()[0]
    tm = pir.process_queries_noalloc(fleet0, fleet1, qd[0], qd[1], got[0], got[1])
    free1 = torch.cuda.mem_get_info()[0]
    torch.cuda.synchronize()
    _say(f"[noalloc] free device memory before / after: {free0} / {free1}")
    host = lambda ts: [bytes(t.cpu().numpy()) if t is not None else None for t in ts]  # noqa: E731
    return host(want), host(got), free0 - free1, tm


def test_noalloc_matches_process_queries_device(pir):
    import torch
    full = _shards(pir, 2048, 17, 700)
    half = _shards(pir, 1024, 8, 800)
    big = [full[k % 17][0] for k in range(129)]   # full secondary group with its conjugate
    small = [g for g, _ in half]
    rng = pir.Rng(bytes([77] * 32))
    # 1: one query on 129 full-shard entries (the largest scratch: l_s = N)
    q, sk = pir.prepare_query(rng, 8 * 2048 - 1, 128, 2048, 2048)
    _say("[noalloc] call 1: 129 full shards, one query")
    want, got, grew, tm = _both(pir, torch, big, None, q, None)
    assert got[0] == want[0]
    assert np.array_equal(pir.process_reply(got[0], sk), full[128 % 17][1][8 * 2048 - 1])
    assert grew == 0
    # 2: smaller, on the same array (8 half shards, both slots: one joint pass)
    q0, sk0 = pir.prepare_query(rng, 3, 7, 1024, 64)
    q1, sk1 = pir.prepare_query(rng, 8 * 1024 - 2, 2, 1024, 64)
    _say("[noalloc] call 2: 8 half shards, two slots")
    want, got, grew, tm = _both(pir, torch, small, small, q0, q1)
    assert got == want
    assert np.array_equal(pir.process_reply(got[0], sk0), half[7][1][3])
    assert np.array_equal(pir.process_reply(got[1], sk1), half[2][1][8 * 1024 - 2])
    assert grew == 0
    # 3: one slot on a fleet of 3 (three-array inner product, l_s = 32)
    q, sk = pir.prepare_query(rng, 100, 1, 1024, 32)
    _say("[noalloc] call 3: 3 half shards, one query")
    want, got, grew, tm = _both(pir, torch, small[:3], None, q, None)
    assert got[0] == want[0] and grew == 0
    assert np.array_equal(pir.process_reply(got[0], sk), half[1][1][100])
    # 4: back to the big fleet after the small ones
    q, sk = pir.prepare_query(rng, 5, 17, 2048, 2048)
    _say("[noalloc] call 4: 129 full shards again")
    want, got, grew, tm = _both(pir, torch, big, None, q, None)
    assert got[0] == want[0] and grew == 0
    assert np.array_equal(pir.process_reply(got[0], sk), full[0][1][5])


def test_noalloc_refuses_what_does_not_fit(pir):
    """both slots on 129 full-shard entries need about 400 MiB of scratch:
    DNSPIR_E_SCRATCH, and the next call that fits still answers exactly"""
    import torch
    from dnspir_amd import DnspirError
    full = _shards(pir, 2048, 2, 900)
    big = [full[k % 2][0] for k in range(129)]
    rng = pir.Rng(bytes([78] * 32))
    q0, _ = pir.prepare_query(rng, 1, 0, 2048, 2048)
    q1, sk1 = pir.prepare_query(rng, 2, 1, 2048, 2048)
    rb = pir.reply_bytes()
    qd0, qd1 = _dev(torch, q0), _dev(torch, q1)
    r0 = torch.zeros(rb, dtype=torch.uint8, device="cuda")
    r1 = torch.zeros(rb, dtype=torch.uint8, device="cuda")
    _say("[noalloc] two slots on 129 full shards: expect DNSPIR_E_SCRATCH")
    with pytest.raises(DnspirError) as ei:
        pir.process_queries_noalloc(big, big, qd0, qd1, r0, r1)
    assert ei.value.code == -15
    assert not r0.any().item() and not r1.any().item()
    pir.process_queries_noalloc(big, None, qd1, None, r1, None)
    want, _ = pir.process_queries(big, None, q1, None)
    assert bytes(r1.cpu().numpy()) == want[0]
    assert np.array_equal(pir.process_reply(want[0], sk1), full[1][1][2])
