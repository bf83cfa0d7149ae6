"""Plain polynomial arithmetic over Z[X]/(X^N + 1) with Python integers.

Used by the tests to state closed forms of the reference algorithms
independently of any NTT: products are schoolbook negacyclic convolutions and
Galois automorphisms act on coefficient indices.
"""
import numpy as np

Q = 67104769 * 33550337


def center(x, m):
    x %= m
    return x - m if x > m // 2 else x


def automorphism(f, g):
    """sigma_g(f)(X) = f(X^g) in Z[X]/(X^N + 1)."""
    N = len(f)
    out = [0] * N
    for i, c in enumerate(f):
        e = (i * g) % (2 * N)
        if e >= N:
            out[e - N] -= c
        else:
            out[e] += c
    return out


def monomial_mul(f, e):
    """f * X^e."""
    N = len(f)
    out = [0] * N
    for i, c in enumerate(f):
        k = (i + e) % (2 * N)
        if k >= N:
            out[k - N] -= c
        else:
            out[k] += c
    return out


def mul(f, g, mod=None):
    N = len(f)
    out = [0] * N
    for i, a in enumerate(f):
        if a == 0:
            continue
        for j, b in enumerate(g):
            k = i + j
            if k >= N:
                out[k - N] -= a * b
            else:
                out[k] += a * b
    if mod is not None:
        out = [x % mod for x in out]
    return out


def add(f, g, mod=None):
    out = [a + b for a, b in zip(f, g)]
    if mod is not None:
        out = [x % mod for x in out]
    return out


def to_int_list(a):
    return [int(x) for x in np.asarray(a).ravel()]
