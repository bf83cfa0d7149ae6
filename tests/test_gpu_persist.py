"""Persisted shards on the GPU: a preprocessed shard written with
dnspir_shard_save and resumed with dnspir_shard_load (no preprocessing)
answers byte for byte as the original; a corrupted, truncated or foreign file
and too little device memory are rejected."""
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
    return p


def _rand_q(seed, n=1):
    return np.random.default_rng(seed).integers(0, orc.Q, (n, N), dtype=np.uint64)


@pytest.mark.parametrize("l", [1024, 2048, 64])
def test_save_load_round_trip(pir, tmp_path, l):
    """The resumed shard answers as an oracle shard preprocessed from the same
    entries: get_a, perform_pir and a reply in which it is the queried shard,
    alone and beside the shard it was saved from."""
    e = np.random.default_rng(40 + l).integers(0, 1025, (8 * l, N)).astype(np.uint16)
    g = pir.PIRDatabase(l)
    g.set_db(e)
    _say(f"[persist] oracle shard l={l}")
    o = orc.Shard(11, l)
    o.set_db(e)
    path = tmp_path / f"s{l}.shard"
    _say(f"[persist] save l={l}")
    g.save(str(path))
    assert path.stat().st_size == 64 + pir.required_bytes(l)
    assert pir.shard_file_info(str(path)) == (l, 0, pir.required_bytes(l))
    del g  # the resumed shard must not depend on the original's device memory
    r = pir.PIRDatabase.load(str(path))
    _say(f"[persist] loaded l={l}")
    assert np.array_equal(r.get_a(), o.get_a())
    qb, gk = _rand_q(1)[0], _rand_q(2, 2)
    qc = _rand_q(3)[0] if l == N else None
    assert np.array_equal(r.perform_pir(qb, qc, gk), o.perform_pir(qb, qc, gk))
    for fleet_len in (1, 2):
        ls = orc.choose_index_group_secondary(fleet_len)
        rng = pir.Rng(bytes([8, fleet_len] * 16))
        q, sk = pir.prepare_query(rng, 8 * l - 1, fleet_len - 1, l, ls)
        other = pir.PIRDatabase(l)
        other.set_db(e[::-1].copy())
        ofl = [orc.Shard(11, l)] if fleet_len == 2 else []
        for x in ofl:
            x.set_db(e[::-1].copy())
        got, _ = pir.process_queries(([other] if fleet_len == 2 else []) + [r], None, q, None)
        want, _, _ = orc.process_queries(ofl + [o], None, q, None)
        assert got == want, fleet_len
        assert np.array_equal(pir.process_reply(got[0], sk), e[8 * l - 1])


def test_damaged_files_are_rejected(pir, tmp_path):
    import torch
    from dnspir_amd import DnspirError
    l = 64
    e = np.random.default_rng(7).integers(0, 1025, (8 * l, N)).astype(np.uint16)
    g = pir.PIRDatabase(l)
    g.set_db(e)
    path = tmp_path / "ok.shard"
    g.save(str(path))
    data = bytearray(path.read_bytes())

    def expect(blob, code, tag):
        f = tmp_path / f"{tag}.shard"
        f.write_bytes(bytes(blob))
        with pytest.raises(DnspirError) as ei:
            pir.PIRDatabase.load(str(f))
        assert ei.value.code == code, tag

    flipped = bytearray(data)
    flipped[64 + len(data) // 2] ^= 0x10
    expect(flipped, -14, "flipped")
    expect(data[:-8], -14, "truncated")
    expect(data + b"\0" * 8, -14, "trailing")
    foreign = bytearray(data)
    foreign[24] ^= 1  # layout fingerprint (bytes 24-31) of another build
    expect(foreign, -14, "foreign")
    small = torch.empty(1024, dtype=torch.uint8, device="cuda")
    with pytest.raises(DnspirError) as ei:
        pir.PIRDatabase.load(str(path), memory=small)
    assert ei.value.code == -6
    with pytest.raises(DnspirError) as ei:
        pir.PIRDatabase(l, memory=small)  # the same check when creating a shard
    assert ei.value.code == -6
    # the intact file still loads
    r = pir.PIRDatabase.load(str(path))
    assert np.array_equal(r.get_a(), g.get_a())
