"""ChaCha block function of the CPU restatement against OpenSSL's ChaCha20.

Pins the quarter round, state layout and feed-forward of oracle/transcript.c (rand_chacha 0.9.0
uses the RFC 8439 core with 8 rounds; the 20-round block runs the same code with rounds = 20).
The ChaCha8Rng buffer/word consumption and seed_from_u64 layout are parity unpinned.
"""
import shutil
import subprocess

import numpy as np
import pytest

pytestmark = pytest.mark.skipif(shutil.which("openssl") is None, reason="openssl not installed")


def _openssl_keystream(key: bytes, counter: int, nonce: bytes, nbytes: int) -> bytes:
    iv = counter.to_bytes(4, "little") + nonce
    r = subprocess.run(["openssl", "enc", "-chacha20", "-K", key.hex(), "-iv", iv.hex()],
                       input=bytes(nbytes), capture_output=True, check=True)
    return r.stdout


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_chacha20_ietf_block_matches_openssl(oracle, seed):
    rng = np.random.default_rng(seed)
    key = bytes(range(32)) if seed == 0 else rng.bytes(32)
    nonce = bytes([0, 0, 0, 9, 0, 0, 0, 0x4A, 0, 0, 0, 0]) if seed == 0 else rng.bytes(12)
    counter = 1 if seed == 0 else int(rng.integers(0, 1 << 31))
    nblk = 4
    ks = _openssl_keystream(key, counter, nonce, 64 * nblk)
    k32 = np.frombuffer(key, dtype="<u4").astype(np.uint32).copy()
    n32 = np.frombuffer(nonce, dtype="<u4").astype(np.uint32).copy()
    for b in range(nblk):
        out = np.zeros(16, np.uint32)
        oracle.lib().or_chacha_ietf_block(k32, counter + b, n32, 20, out)
        assert out.astype("<u4").tobytes() == ks[64 * b:64 * (b + 1)]


def test_chacha_block_64bit_counter_layout(oracle):
    # or_chacha_block: words 12-13 = 64-bit block counter, 14-15 = 64-bit stream id
    rng = np.random.default_rng(5)
    k32 = rng.integers(0, 1 << 32, size=8, dtype=np.uint64).astype(np.uint32)
    for ctr, stream in [(0, 0), (7, 0), ((3 << 32) | 5, 0x1122334455667788)]:
        a = np.zeros(16, np.uint32)
        oracle.lib().or_chacha_block(k32, ctr, stream, 20, a)
        nonce = np.array([ctr >> 32, stream & 0xFFFFFFFF, stream >> 32], dtype=np.uint32)
        b = np.zeros(16, np.uint32)
        oracle.lib().or_chacha_ietf_block(k32, ctr & 0xFFFFFFFF, nonce, 20, b)
        assert np.array_equal(a, b)
        ks = _openssl_keystream(k32.astype("<u4").tobytes(), ctr & 0xFFFFFFFF,
                                nonce.astype("<u4").tobytes(), 64)
        assert a.astype("<u4").tobytes() == ks


def test_eight_rounds_differ_from_twenty(oracle):
    k32 = np.arange(8, dtype=np.uint32)
    a = np.zeros(16, np.uint32)
    b = np.zeros(16, np.uint32)
    oracle.lib().or_chacha_block(k32, 0, 0, 8, a)
    oracle.lib().or_chacha_block(k32, 0, 0, 20, b)
    assert not np.array_equal(a, b)
