"""Known-answer tests of set_transformed_entries at N = 8
(base_pir.rs:1281-1411, restated) and closed forms of K5 / K6 / K1 checked
with plain polynomial arithmetic (CPU only)."""
import numpy as np
import pytest

import polyref as pr
import pyoracle as orc

T = orc.T


def _kat_check(shard, expected_rows, l):
    """TE row k, lane 0, modulus p1 equals NTT(expected_k * l^-1) with slot i
    evaluated at zeta^-(2i+1) (base_pir.rs:1298)."""
    te, zeta = shard.export_te_p1()
    p = orc.P1
    N = shard.N
    factor = pow(l, -1, T)
    zinv = pow(zeta, p - 2, p)
    assert len(expected_rows) == shard.L
    for k, row in enumerate(expected_rows):
        coeffs = [pr.center(factor * x, T) % p for x in row]
        for i in range(N):
            want = sum(coeffs[j] * pow(zinv, (2 * i + 1) * j, p) for j in range(N)) % p
            assert int(te[k, 0, i]) == want, (k, i)


def _entries(n, N, which, value):
    e = np.zeros((n, N), np.uint16)
    if value == "one":
        e[which, 0] = 1
    else:  # X
        e[which, 1] = 1
    return e


Z = [0] * 8


def _row(pos, val):
    r = [0] * 8
    r[pos] = val % T
    return r


@pytest.mark.parametrize("l", [8, 4])
def test_kat_full_group_and_full_g_subgroup(l):
    N = 8
    sh = orc.Shard(3, l)
    sh.set_transformed_entries(_entries(8 * l, N, 0, "one"))
    _kat_check(sh, [_row(0, 1)] * 4, l)
    sh.set_transformed_entries(_entries(8 * l, N, 8, "one"))
    _kat_check(sh, [_row(7, -1)] * 4, l)
    sh.set_transformed_entries(_entries(8 * l, N, 0, "X"))
    _kat_check(sh, [_row(5, 1), _row(1, -1), _row(5, -1), _row(1, 1)], l)


def test_kat_small_subgroup():
    N, l = 8, 2
    sh = orc.Shard(3, l)
    sh.set_transformed_entries(_entries(16, N, 0, "one"))
    _kat_check(sh, [_row(0, 1)] * 2, l)
    sh.set_transformed_entries(_entries(16, N, 8, "one"))
    _kat_check(sh, [_row(6, -1)] * 2, l)
    sh.set_transformed_entries(_entries(16, N, 0, "X"))
    _kat_check(sh, [_row(5, -1), _row(1, 1)], l)


def _closed_form_te(entries, N, l, lane):
    """TE_c = l^-1 sum_j sigma_{5^-(L-1-c)}(a_j) X^(-kj) (main rows) and with
    X^(+kj) (conjugate rows); a_j = entry 8j + lane."""
    L = l if l < N else N // 2
    k = N // 2 // L
    inv5 = pow(5, -1, 2 * N)
    linv = pow(l, -1, T)
    rows = []
    for stream in range(2 if l == N else 1):
        for c in range(L):
            g = pow(inv5, L - 1 - c, 2 * N)
            acc = [0] * N
            for j in range(l):
                a = [int(x) for x in entries[8 * j + lane]]
                term = pr.monomial_mul(pr.automorphism(a, g), (-k * j if stream == 0 else k * j) % (2 * N))
                acc = pr.add(acc, term)
            rows.append([pr.center(linv * x, T) for x in acc])
    return rows


@pytest.mark.parametrize("logN,l", [(4, 16), (4, 8), (4, 2), (5, 32), (5, 16), (5, 4)])
def test_k5_closed_form(logN, l):
    N = 1 << logN
    rng = np.random.default_rng(logN * 100 + l)
    entries = rng.integers(0, T, (8 * l, N)).astype(np.uint16)
    sh = orc.Shard(logN, l)
    sh.set_transformed_entries(entries)
    for lane in (0, 5):
        want = _closed_form_te(entries, N, l, lane)
        got = []
        for stream in range(2 if sh.conj else 1):
            for c in range(sh.L):
                got.append([int(x) for x in sh.row_coeffs(0, stream, c, lane)])
        assert got == want


def _decomp(x):
    """balanced base-2^26 digits of the centered lift (base_pir.rs:544-570)"""
    x = pr.center(x, orc.Q)
    hi = (x + (1 << 25)) >> 26
    lo = x - (hi << 26)
    hi2 = (hi + (1 << 25)) >> 26
    lo2 = hi - (hi2 << 26)
    return lo, lo2


@pytest.mark.parametrize("logN,l,seeds", [(4, 16, 0), (4, 4, 1), (5, 16, 0)])
def test_k6_and_k1_closed_form(logN, l, seeds):
    """preprocess (a-chain + key-switch digits) and perform_pir against ring
    arithmetic: a_{i+1} = D0(s(a_i)) gk_a0 + D1(s(a_i)) gk_a1 + qa TE_i,
    b = sum_streams sum_i s^(L-1-i)(D0_i gkb0 + D1_i gkb1 + TE_i qb)."""
    N = 1 << logN
    Q = orc.Q
    rng = np.random.default_rng(7 * logN + l)
    entries = rng.integers(0, T, (8 * l, N)).astype(np.uint16)
    sh = orc.Shard(logN, l, used_seeds=seeds)
    sh.set_db(entries)
    gka = [pr.to_int_list(orc.expand(logN, orc.seed(d))) for d in range(2)]
    nst = 2 if sh.conj else 1
    qa = [pr.to_int_list(orc.expand(logN, orc.seed((2 if seeds == 0 else 4) + s))) for s in range(nst)]
    lanes = (0, 3)
    reply_a = {ln: [0] * N for ln in lanes}
    for lane in lanes:
        for st in range(nst):
            a = [0] * N
            for i in range(sh.L):
                ga = pr.automorphism(a, 5)
                d = [_decomp(x) for x in ga]
                d0 = [x[0] for x in d]
                d1 = [x[1] for x in d]
                assert [int(v) for v in sh.row_coeffs(1, st, i, lane)] == d0
                assert [int(v) for v in sh.row_coeffs(2, st, i, lane)] == d1
                te = [int(v) for v in sh.row_coeffs(0, st, i, lane)]
                a = pr.add(pr.add(pr.mul(d0, gka[0]), pr.mul(d1, gka[1])), pr.mul(qa[st], te), Q)
            reply_a[lane] = pr.add(reply_a[lane], a, Q)
    got_a = sh.get_a()
    for lane in lanes:
        assert [int(x) for x in got_a[lane]] == reply_a[lane]
    # K1: perform_pir with random query vectors
    qb = rng.integers(0, Q, N, dtype=np.uint64)
    qbc = rng.integers(0, Q, N, dtype=np.uint64) if sh.conj else None
    gkb = rng.integers(0, Q, (2, N), dtype=np.uint64)
    out = sh.perform_pir(qb, qbc, gkb)
    qs = [pr.to_int_list(qb)] + ([pr.to_int_list(qbc)] if sh.conj else [])
    gk = [pr.to_int_list(gkb[0]), pr.to_int_list(gkb[1])]
    for lane in lanes:
        b = [0] * N
        for st in range(nst):
            for i in range(sh.L):
                d0 = [int(v) for v in sh.row_coeffs(1, st, i, lane)]
                d1 = [int(v) for v in sh.row_coeffs(2, st, i, lane)]
                te = [int(v) for v in sh.row_coeffs(0, st, i, lane)]
                w = pr.add(pr.add(pr.mul(d0, gk[0]), pr.mul(d1, gk[1])), pr.mul(te, qs[st]), Q)
                g = pow(5, sh.L - 1 - i, 2 * N)
                b = pr.add(b, pr.automorphism(w, g), Q)
        assert [int(x) for x in out[lane]] == b
