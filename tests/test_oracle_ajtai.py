"""Balanced base-2 decomposition, DEC digit split and the Ajtai commitment of the CPU restatement.

Pins oracle/ajtai.c and oracle/shard.c helpers against properties fixed by the reference:
  * decomp_b_row_major, b = 2 balanced (neo-ajtai/src/decomp.rs:118-158): digits in {-1, 0, 1},
    a non-zero digit carries the sign of the remaining value, sum_rho 2^rho Z[rho] = z for
    |z| < 2^54, and beyond that the 54 digits keep z mod 2^54 (truncation);
  * split_b with b = 2 (neo-reductions/src/common.rs:20-68, balanced_divrem): same rule, 12 digits,
    values outside [-4095, 4095] rejected;
  * commit (neo-ajtai/src/commit.rs): c_i = sum_j cf(a_ij * zhat_j mod Phi_81), checked by a
    big-integer brute force on small m, plus linearity in Z.
"""
import numpy as np
import pytest

from kmath import Q, D


def _phi81_reduce(p):
    p = [c % Q for c in p]
    for i in range(len(p) - 1, D - 1, -1):
        c = p[i]
        if c:
            p[i] = 0
            p[i - D] = (p[i - D] - c) % Q
            p[i - 27] = (p[i - 27] - c) % Q
    return p[:D]


def _bal(f):
    return f if f <= (Q - 1) // 2 else f - Q


def _check_digit_rule(vals, digits):
    for c, v in enumerate(vals):
        a = v
        for rho in range(digits.shape[0]):
            d = int(digits[rho, c])
            assert d in (-1, 0, 1)
            assert (d == 0) == (a % 2 == 0)
            if d:
                assert (d > 0) == (a > 0)
            a = (a - d) // 2


def test_decomp_round_trip_in_range(oracle):
    rng = np.random.default_rng(21)
    m = 512
    lim = (1 << 54) - 1
    vals = [int(v) for v in rng.integers(-lim, lim + 1, size=m - 6)] + [0, 1, -1, lim, -lim, 2]
    z = np.array([v % Q for v in vals], dtype=np.uint64)
    Z = oracle.decomp(z)
    assert Z.shape == (D, m)
    _check_digit_rule(vals, Z)
    for c, v in enumerate(vals):
        assert sum(int(Z[rho, c]) << rho for rho in range(D)) == v


def test_decomp_truncates_beyond_54_digits(oracle):
    rng = np.random.default_rng(22)
    m = 256
    z = rng.integers(0, Q, size=m, dtype=np.uint64)
    Z = oracle.decomp(z)
    for c in range(m):
        v = _bal(int(z[c]))
        s = sum(int(Z[rho, c]) << rho for rho in range(D))
        assert abs(s) < (1 << 54)
        assert (s - v) % (1 << 54) == 0


def test_split_b_round_trip_and_flags(oracle):
    rng = np.random.default_rng(23)
    vals = np.concatenate([rng.integers(-4095, 4096, size=D * 64 - 4), [4095, -4095, 0, 0]]).astype(np.int32)
    vals[D * 10:D * 11] = 0
    outs, nz = oracle.split_b(vals, 12)
    digits = np.stack(outs).reshape(12, -1)
    _check_digit_rule([int(v) for v in vals], digits)
    rec = sum(digits[i].astype(np.int64) << i for i in range(12))
    assert np.array_equal(rec, vals.astype(np.int64))
    assert nz == [bool(np.any(digits[i])) for i in range(12)]


@pytest.mark.parametrize("bad", [4096, -4096, 1 << 20])
def test_split_b_rejects_out_of_range(oracle, bad):
    vals = np.zeros(D * 8, np.int32)
    vals[17] = bad
    with pytest.raises(ValueError):
        oracle.split_b(vals, 12)


def _commit_bruteforce(pp, kappa, m, Z):
    out = []
    for i in range(kappa):
        acc = [0] * (2 * D - 1)
        for j in range(m):
            a = [int(x) for x in pp[(i * m + j) * D:(i * m + j + 1) * D]]
            zh = [int(Z[rho, j]) for rho in range(D)]
            for u in range(D):
                if a[u]:
                    for v in range(D):
                        if zh[v]:
                            acc[u + v] += a[u] * zh[v]
        out.extend(_phi81_reduce(acc))
    return out


@pytest.mark.parametrize("kappa,m", [(2, 4), (3, 8)])
def test_commit_matches_bruteforce(oracle, kappa, m):
    rng = np.random.default_rng(24 + m)
    pp = oracle.ajtai_setup(bytes(range(32)), kappa, m)
    assert pp.shape == (kappa * m * D,) and int(pp.max()) < Q
    Z = rng.integers(-1, 2, size=(D, m)).astype(np.int8)
    c = oracle.ajtai_commit(pp, kappa, m, Z)
    assert [int(v) for v in c] == _commit_bruteforce(pp, kappa, m, Z)


def test_commit_is_linear(oracle):
    # commitments take balanced base-2 digits; disjoint supports keep Z1 + Z2 in {-1, 0, 1}
    rng = np.random.default_rng(25)
    kappa, m = 16, 64
    pp = oracle.ajtai_setup(b"\x01" * 32, kappa, m)
    Z = rng.integers(-1, 2, size=(D, m)).astype(np.int8)
    mask = rng.integers(0, 2, size=(D, m)).astype(bool)
    Z1 = np.where(mask, Z, 0).astype(np.int8)
    Z2 = np.where(mask, 0, Z).astype(np.int8)
    c = oracle.ajtai_commit(pp, kappa, m, Z)
    c1 = oracle.ajtai_commit(pp, kappa, m, Z1)
    c2 = oracle.ajtai_commit(pp, kappa, m, Z2)
    assert [int(v) for v in c] == [(int(a) + int(b)) % Q for a, b in zip(c1, c2)]
    cn = oracle.ajtai_commit(pp, kappa, m, (-Z).astype(np.int8))
    assert [int(v) for v in cn] == [(-int(a)) % Q for a in c]
    assert not np.any(oracle.ajtai_commit(pp, kappa, m, np.zeros((D, m), np.int8)))


def test_pp_setup_is_seed_deterministic(oracle):
    a = oracle.ajtai_setup(b"\x02" * 32, 16, 64)
    b = oracle.ajtai_setup(b"\x02" * 32, 16, 64)
    c = oracle.ajtai_setup(b"\x03" * 32, 16, 64)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert int(a.max()) < Q
