"""Shared CPU-bus helpers of the CPU restatement (neo-fold/src/memory_sidecar/cpu_bus.rs).

decode_vector is checked against Python big integers; each bus time column against a direct
gather of the decoded witness at bus_cell(col, j) = bus_base + col * chunk + j (bus_layout.rs);
each opening row y[col][rho] against the multilinear extension of that digit row's time vector
(nonzero only at t = m_in + j) evaluated at r, and y_scalar against the multilinear extension of
the decoded time column returned by build_time_sparse_from_bus_col."""
import numpy as np
import pytest

from kmath import D, K0, Q, kadd, kscale, mle, rand_k


def _layout(rng, m=1024, m_in=1, chunk=64, bus_cols=7):
    Z = rng.integers(-1, 2, size=(D, m)).astype(np.int8)
    bus_base = m - bus_cols * chunk
    return Z, m, m_in, bus_base, chunk, bus_cols


def test_decode(oracle):
    rng = np.random.default_rng(1)
    Z = rng.integers(-1, 2, size=(D, 300)).astype(np.int8)
    z = oracle.bus_decode(Z)
    for c in range(300):
        assert int(z[c]) == sum(int(Z[r, c]) << r for r in range(D)) % Q


@pytest.mark.parametrize("steps_len", [64, 40])
def test_time_columns_and_openings(oracle, steps_len):
    rng = np.random.default_rng(2 + steps_len)
    Z, m, m_in, bus_base, chunk, bus_cols = _layout(rng)
    ell_n = 7  # m_in + chunk <= 2^ell_n
    z = oracle.bus_decode(Z)
    r = [rand_k(rng) for _ in range(ell_n)]
    y, ys = oracle.bus_openings(Z, m_in, bus_base, chunk, bus_cols, r)
    for col in range(bus_cols):
        idx, vals = oracle.bus_time_col(z, m_in, bus_base, chunk, bus_cols, col, steps_len, 1 << ell_n)
        exp = [(m_in + j, int(z[bus_base + col * chunk + j])) for j in range(steps_len)
               if int(z[bus_base + col * chunk + j]) != 0]
        assert [int(i) for i in idx] == [t for t, _ in exp]
        assert vals == [(v, 0) for _, v in exp]
        for rho in range(D):
            tv = [0] * (1 << ell_n)
            for j in range(chunk):
                tv[m_in + j] = int(Z[rho, bus_base + col * chunk + j]) % Q
            assert y[col][rho] == mle(tv, r)
        assert all(v == K0 for v in y[col][D:])
        if steps_len == chunk:
            dense = [0] * (1 << ell_n)
            for t, v in zip(idx, vals):
                dense[int(t)] = v[0]
            assert ys[col] == mle(dense, r)
        s = K0
        for rho in range(D):
            s = kadd(s, kscale(y[col][rho], pow(2, rho, Q)))
        assert ys[col] == s


def test_rejects_bad_layout(oracle):
    rng = np.random.default_rng(9)
    Z, m, m_in, bus_base, chunk, bus_cols = _layout(rng)
    with pytest.raises(ValueError):
        oracle.bus_openings(Z, m_in, bus_base, chunk, bus_cols, [rand_k(rng) for _ in range(6)])  # 2^6 < m_in+chunk
    z = oracle.bus_decode(Z)
    with pytest.raises(ValueError):
        oracle.bus_time_col(z, m_in, bus_base, chunk, bus_cols, bus_cols, chunk, 128)
