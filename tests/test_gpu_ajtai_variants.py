"""The tile shapes of the commitment GEMM (16 x 16 x 64 MFMAs in 192 x 352 blocks, 32 x 32 x 32
MFMAs in 256 x 256 blocks with direct-to-LDS staging of int8 digits, and the same blocks on digits
packed to 2-bit codes) give the same commitments at every shape the
prover uses, each checked against the oracle: the shard's DEC children (12 at 2^16), a padded
N tile (13 and 5 witnesses), the sweep's 64 witnesses, a short m, and the K-step-major digit path
(2^18).  Witnesses are dense where the oracle is cheap, else digits in 64 random columns (a
commitment is a sum over columns)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
D = 54


@pytest.fixture
def forced(ns):
    yield ns.ajtai_force_variant
    ns.ajtai_force_variant(0)


@pytest.mark.parametrize("logm,nw", [(10, 3), (16, 12), (16, 13), (16, 64), (18, 5)])
def test_tile_variants_agree_with_oracle(ns, oracle, forced, logm, nw):
    print(f"commit tile variants m=2^{logm} nw={nw}", flush=True)
    rng = np.random.default_rng(1000 + logm * 7 + nw)
    m = 1 << logm
    pp = ns.PublicParams(m, seed=bytes([11] * 32), keep_coeffs=True)
    ppc = pp.coeffs()
    ppc3 = ppc.reshape(16, m, D)
    Zs, cols = [], []
    for w in range(nw):
        if logm <= 10 or w in (0, nw - 1):
            Zs.append(rng.integers(-1, 2, size=(D, m), dtype=np.int8))
            cols.append(None)
        else:
            c = np.sort(rng.choice(m, size=64, replace=False))
            z = np.zeros((D, m), np.int8)
            z[:, c] = rng.integers(-1, 2, size=(D, 64), dtype=np.int8)
            Zs.append(z)
            cols.append(c)
    bufs = [ns.DeviceBuffer.from_array(z.reshape(-1)) for z in Zs]
    got = {}
    for v in (1, 2, 3):
        forced(v)
        got[v] = pp.commit_device(bufs)
    for v in (2, 3):
        assert np.array_equal(got[1], got[v]), f"tile shape {v}"
    check = [0, nw - 1] + list(rng.choice(nw, size=min(4, nw), replace=False))
    for w in sorted(set(int(x) for x in check)):
        if cols[w] is None:
            if logm > 16:
                continue  # dense oracle commitments at 2^18 take minutes; the sparse ones cover it
            want = oracle.ajtai_commit(ppc, 16, m, Zs[w])
        else:
            sub = np.ascontiguousarray(ppc3[:, cols[w], :]).reshape(-1)
            want = oracle.ajtai_commit(sub, 16, 64, np.ascontiguousarray(Zs[w][:, cols[w]]))
        assert np.array_equal(got[2][w], want), f"witness {w}"
