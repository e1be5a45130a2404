"""The product's host Poseidon2 (transcript width 8, CCS-digest width 16) against the oracle's
independent restatement, on the scalar and on the dispatched (AVX-512 where present) path.
No GPU needed: the permutation is host code of libnightstream.so."""
import numpy as np
import pytest

import nightstream_amd as ns

Q = (1 << 64) - (1 << 32) + 1


def states(width, n, seed):
    rng = np.random.default_rng(seed)
    out = [np.zeros(width, np.uint64), np.full(width, Q - 1, np.uint64)]
    edge = np.array([0, 1, Q - 1, Q - 2, (1 << 32) - 1, 1 << 32, (1 << 63), Q - (1 << 32)], np.uint64)
    out.append(np.resize(edge, width))
    for _ in range(n):
        out.append((rng.integers(0, 1 << 63, size=width, dtype=np.uint64) * 2 + rng.integers(0, 2, size=width, dtype=np.uint64)) % np.uint64(Q))
    return out


@pytest.mark.parametrize("force_scalar", [True, False])
def test_transcript_permutation_matches_oracle(oracle, force_scalar):
    for st in states(8, 200, 7):
        want = oracle.p2_permute(st)
        got = ns.poseidon2_permute(st, force_scalar=force_scalar)
        assert np.array_equal(got, want)


@pytest.mark.parametrize("force_scalar", [True, False])
def test_digest_permutation_matches_oracle(oracle, force_scalar):
    for st in states(16, 100, 8):
        want = oracle.p2_permute(st)
        got = ns.poseidon2_permute(st, force_scalar=force_scalar)
        assert np.array_equal(got, want)


def test_permutation_rejects_bad_input():
    with pytest.raises(Exception):
        ns.poseidon2_permute(np.zeros(12, np.uint64))
    with pytest.raises(Exception):
        ns.poseidon2_permute(np.full(8, Q, np.uint64))
