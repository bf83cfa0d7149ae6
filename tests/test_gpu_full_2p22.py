"""The 2^22 configuration at full size: 256 distinct full shards resident
(153 GiB of slabs), the fleet the bench times. The oracle cannot preprocess
256 full shards within a test, so the checks here are properties that do not
depend on size:
- every reply decrypts to its entry;
- the two-array inner product (IPGK0 folded into per-query weights) and the
  three-array one (all stored arrays), two different kernels, give identical
  bytes;
- a query's reply is the same alone, repeated, and as one slot of a two-query
  call (which runs the two-query form of the two-array kernel).
Byte parity with the oracle on distinct shards is in test_gpu_distinct.py (17
to 20 shards) and on the 2^22 shape in test_gpu_configs.py."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

N = 2048
L_P = 2048
SHARDS = 256
SEED0 = 7000


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def _entries(torch, gen, k):
    """Shard k's entries: seeded uniform coefficients below 1025, on the device."""
    gen.manual_seed(SEED0 + k)
    return torch.randint(0, 1025, (8 * L_P, N), dtype=torch.int16, device="cuda", generator=gen)


def test_2p22_full_size_distinct_shards():
    import torch
    from dnspir_amd import pir

    pir.init(0)
    assert pir.get_database_shape(SHARDS * 8 * L_P) == (L_P, L_P)
    gen = torch.Generator(device="cuda")
    _say(f"[full 2^22] preprocessing {SHARDS} distinct full shards")
    shards = []
    for k in range(SHARDS):
        c = _entries(torch, gen, k)
        s = pir.PIRDatabase(L_P)
        s.set_db_device(c)
        shards.append(s)
        del c
    torch.cuda.synchronize()
    try:
        rng = pir.Rng(bytes([21] + [0] * 31))
        # (entry within the shard, shard): the first and last entries of the
        # database and two in between
        picks = [(0, 0), (5000, 97), (12345, 200), (8 * L_P - 1, SHARDS - 1)]
        qs = [pir.prepare_query(rng, i, j, L_P, L_P) for i, j in picks]
        three = pir.Context()
        three.set_k1_mode(3)
        two = pir.Context()
        two.set_k1_mode(2)
        replies = []
        for (i, j), (q, sk) in zip(picks, qs):
            _say(f"[full 2^22] query entry {i} of shard {j}: two arrays, three arrays, repeated")
            r2, _ = pir.process_queries(shards, None, q, None, ctx=two)
            r3, _ = pir.process_queries(shards, None, q, None, ctx=three)
            assert r2[0] == r3[0], (i, j)
            again, _ = pir.process_queries(shards, None, q, None, ctx=two)
            assert again[0] == r2[0], (i, j)
            want = _entries(torch, gen, j)[i].cpu().numpy().astype(np.uint16)
            assert np.array_equal(pir.process_reply(r2[0], sk), want), (i, j)
            replies.append(r2[0])
        _say("[full 2^22] two queries in one call (both slots on the same fleet)")
        for a, b in ((0, 3), (1, 2)):
            joint, _ = pir.process_queries(shards, shards, qs[a][0], qs[b][0], ctx=two)
            assert joint[0] == replies[a] and joint[1] == replies[b], (a, b)
    finally:
        del shards
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
