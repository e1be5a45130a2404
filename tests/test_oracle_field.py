"""F_q, K and R_q arithmetic of the CPU restatement against Python big integers.

Pins oracle/field.c: neo-math/src/field.rs:10-89 (Goldilocks, K = F_q[u]/(u^2 - 7)),
ring.rs:54-66,187-203 (product mod Phi_81 = X^54 + X^27 + 1), commit.rs:46-54 (rot_step).
"""
import numpy as np
import pytest

from kmath import Q, D, kinv, kmul

EDGE = [0, 1, 2, Q - 1, Q - 2, (1 << 32) - 1, 1 << 32, (1 << 63), Q >> 1, (Q >> 1) + 1, 0xFFFFFFFF00000000]


def _rand_fq(rng, n):
    v = rng.integers(0, Q, size=n, dtype=np.uint64)
    return np.concatenate([np.array(EDGE, dtype=np.uint64), v])


def test_fq_mul_matches_bigint(oracle):
    rng = np.random.default_rng(11)
    a = _rand_fq(rng, 4000)
    b = np.roll(_rand_fq(rng, 4000), 3)
    out = np.zeros_like(a)
    oracle.lib().or_fq_mul_batch(a, b, out, len(a))
    exp = [(int(x) * int(y)) % Q for x, y in zip(a, b)]
    assert [int(v) for v in out] == exp


def test_k_mul_and_inverse(oracle):
    rng = np.random.default_rng(12)
    n = 2000
    a = rng.integers(0, Q, size=2 * n, dtype=np.uint64)
    b = rng.integers(0, Q, size=2 * n, dtype=np.uint64)
    a[:2] = [Q - 1, Q - 1]
    b[:2] = [Q - 1, 1]
    out = np.zeros_like(a)
    oracle.lib().or_k_mul_batch(a, b, out, n)
    for i in range(n):
        x, y = (int(a[2 * i]), int(a[2 * i + 1])), (int(b[2 * i]), int(b[2 * i + 1]))
        assert (int(out[2 * i]), int(out[2 * i + 1])) == kmul(x, y)
    inv = np.zeros_like(a)
    oracle.lib().or_k_inv_batch(a, inv, n)
    for i in range(n):
        x = (int(a[2 * i]), int(a[2 * i + 1]))
        xi = (int(inv[2 * i]), int(inv[2 * i + 1]))
        assert xi == kinv(x)
        assert kmul(x, xi) == (1, 0)


def test_seven_is_a_nonresidue():
    # K = F_q[u]/(u^2 - 7) is a field iff 7 is not a square mod q (neo-math/tests/field_ext.rs:19)
    assert pow(7, (Q - 1) // 2, Q) == Q - 1


def _phi81_reduce(p):
    """Long division by X^54 + X^27 + 1 (coefficients mod q)."""
    p = [c % Q for c in p]
    for i in range(len(p) - 1, D - 1, -1):
        c = p[i]
        if c:
            p[i] = 0
            p[i - D] = (p[i - D] - c) % Q
            p[i - 27] = (p[i - 27] - c) % Q
    return p[:D]


def test_rq_mul_matches_long_division(oracle):
    rng = np.random.default_rng(13)
    for trial in range(40):
        if trial % 2:
            a = rng.integers(0, Q, size=D, dtype=np.uint64)
        else:
            a = np.array([int(v) % Q for v in rng.integers(-2, 3, size=D)], dtype=np.uint64)
        b = rng.integers(0, Q, size=D, dtype=np.uint64)
        out = np.zeros(D, np.uint64)
        oracle.lib().or_rq_mul(a, b, out)
        prod = [0] * (2 * D - 1)
        for i in range(D):
            for j in range(D):
                prod[i + j] += int(a[i]) * int(b[j])
        assert [int(v) for v in out] == _phi81_reduce(prod)


def test_x81_is_one_in_rq(oracle):
    # Phi_81 divides X^81 - 1: multiplying by X 81 times is the identity
    rng = np.random.default_rng(14)
    cur = rng.integers(0, Q, size=D, dtype=np.uint64)
    start = cur.copy()
    nxt = np.zeros(D, np.uint64)
    for step in range(81):
        oracle.lib().or_rot_step(cur, nxt)
        cur, nxt = nxt, cur
        if step < 80:
            assert not np.array_equal(cur, start)
    assert np.array_equal(cur, start)


def test_rot_step_is_multiplication_by_x(oracle):
    rng = np.random.default_rng(15)
    x = np.zeros(D, np.uint64)
    x[1] = 1
    for _ in range(10):
        a = rng.integers(0, Q, size=D, dtype=np.uint64)
        nxt = np.zeros(D, np.uint64)
        oracle.lib().or_rot_step(a, nxt)
        viam = np.zeros(D, np.uint64)
        oracle.lib().or_rq_mul(a, x, viam)
        assert np.array_equal(nxt, viam)
        prod = [0] * (2 * D - 1)
        for i in range(D):
            prod[i + 1] += int(a[i])
        assert [int(v) for v in nxt] == _phi81_reduce(prod)


@pytest.mark.parametrize("v", [0, 1, -1, 5, -5, (1 << 62), -(1 << 62), (1 << 63) - 1, -(1 << 63)])
def test_from_i64_and_balanced(oracle, v):
    import ctypes as C
    L = oracle.lib()
    L.or_fq_from_i64.restype = C.c_uint64
    L.or_fq_from_i64.argtypes = [C.c_int64]
    L.or_fq_to_balanced.restype = C.c_int64
    L.or_fq_to_balanced.argtypes = [C.c_uint64]
    f = L.or_fq_from_i64(v)
    assert f == v % Q
    bal = L.or_fq_to_balanced(f)
    assert bal % Q == v % Q and -(Q - 1) // 2 <= bal <= (Q - 1) // 2
