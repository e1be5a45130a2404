"""Python big-integer arithmetic for F_q, K = F_q[u]/(u^2 - 7) and multilinear tables.

Written from the definitions (neo-math/src/field.rs:10-89) for use as an independent check of
the CPU restatement; nothing here calls the library under test.
"""
Q = 0xFFFFFFFF00000001
D = 54
K0, K1 = (0, 0), (1, 0)


def kadd(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def ksub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def kmul(a, b):
    return ((a[0] * b[0] + 7 * a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def kscale(a, s):
    return (a[0] * s % Q, a[1] * s % Q)


def kpow(a, e):
    r = K1
    while e:
        if e & 1:
            r = kmul(r, a)
        a = kmul(a, a)
        e >>= 1
    return r


def kinv(a):
    n = (a[0] * a[0] - 7 * a[1] * a[1]) % Q
    ni = pow(n, Q - 2, Q)
    return (a[0] * ni % Q, (-a[1]) * ni % Q)


def ksum(xs):
    r = K0
    for x in xs:
        r = kadd(r, x)
    return r


def chi(point):
    """chi_point[idx] = prod_b (bit_b(idx) ? point[b] : 1 - point[b]), bit 0 = least significant."""
    w = [K1]
    for p in point:
        nb = ksub(K1, p)
        w = [kmul(v, nb) for v in w] + [kmul(v, p) for v in w]
    return w


def eq_points(p, q):
    acc = K1
    for a, b in zip(p, q):
        acc = kmul(acc, kadd(kmul(ksub(K1, a), ksub(K1, b)), kmul(a, b)))
    return acc


def mle(table, point):
    """Multilinear extension of `table` (values in K or ints) at `point`."""
    w = chi(point)
    acc = K0
    for v, c in zip(table, w):
        if isinstance(v, tuple):
            acc = kadd(acc, kmul(v, c))
        elif v:
            acc = kadd(acc, kscale(c, v % Q))
    return acc


def poly_eval(coeffs, x):
    r = K0
    for c in reversed(coeffs):
        r = kadd(kmul(r, x), c)
    return r


def rand_k(rng):
    return (int(rng.integers(0, Q, dtype="u8")), int(rng.integers(0, Q, dtype="u8")))


def to_arr(ks):
    import numpy as np
    return np.array([v for k in ks for v in k], dtype=np.uint64)


def from_arr(a):
    a = [int(v) for v in a.reshape(-1)]
    return [(a[2 * i], a[2 * i + 1]) for i in range(len(a) // 2)]
