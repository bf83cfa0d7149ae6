"""The benchmark configurations at full size: the 2^22 fleet of 256 distinct
full shards (153 GiB of slabs) the bench times, and 262144 entries as 32
distinct half shards with a 64-query batch. The oracle cannot preprocess 256
full shards within a test, so the checks here are properties that do not
depend on size:
- every reply decrypts to its entry;
- two different kernels give identical bytes: the two-array inner product
  (IPGK0 folded into per-query weights) and the three-array one at 2^22, the
  batched int8-MFMA one and the single-query one at 262144;
- a query's reply is the same alone, repeated, and as one slot of a two-query
  call (which runs the two-query form of the two-array kernel).
Byte parity with the oracle on distinct shards is in test_gpu_distinct.py (17
to 20 shards) and on both shapes in test_gpu_configs.py."""
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


def test_262144_full_size_batch_of_64():
    """262144 entries at full size: 32 distinct half shards (9.6 GiB), one batch
    of 64 clients with their own keys (the int8-MFMA inner product over all
    three arrays). Every reply equals the query's single reply (the two-array
    inner product, pinned to the oracle on distinct shards) and decrypts to its
    entry."""
    import torch
    from dnspir_amd import pir

    pir.init(0)
    lp, ls = pir.get_database_shape(262144)
    n = pir.num_databases(262144)
    assert (lp, ls, n) == (1024, 256, 32)
    gen = torch.Generator(device="cuda")

    def entries(k):
        gen.manual_seed(SEED0 + 1000 + k)
        return torch.randint(0, 1025, (8 * lp, N), dtype=torch.int16, device="cuda", generator=gen)

    _say(f"[full 262144] preprocessing {n} distinct half shards")
    shards = []
    for k in range(n):
        c = entries(k)
        s = pir.PIRDatabase(lp)
        s.set_db_device(c)
        shards.append(s)
        del c
    torch.cuda.synchronize()
    try:
        rng = pir.Rng(bytes([22] + [0] * 31))
        idx = np.random.default_rng(22)
        picks = [(int(idx.integers(0, 8 * lp)), int(idx.integers(0, n))) for _ in range(62)]
        picks += [(0, 0), (8 * lp - 1, n - 1)]
        qs = [pir.prepare_query(rng, i, j, lp, ls) for i, j in picks]
        ctx = pir.Context()
        _say("[full 262144] batch of 64")
        reps, _ = pir.process_batch(shards, [q for q, _ in qs], ctx=ctx)
        single = pir.Context()
        for t, ((i, j), (q, sk)) in enumerate(zip(picks, qs)):
            one, _ = pir.process_queries(shards, None, q, None, ctx=single)
            assert reps[t] == one[0], (t, i, j)
            want = entries(j)[i].cpu().numpy().astype(np.uint16)
            assert np.array_equal(pir.process_reply(reps[t], sk), want), (t, i, j)
    finally:
        del shards
        torch.cuda.synchronize()
        torch.cuda.empty_cache()
