"""The memory sidecar on the GPU at the full chunk size (2^16 cycles, the largest a batch or served
job carries): all 30 device oracles, driven by the grouped batched driver, run valid sum-check
chains (a size-independent property; at small sizes the chunk's bytes are compared with the CPU
restatement in test_gpu_batch.py)."""
import pytest

from test_sidecar import at, kadd

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ell", [13, 16])
def test_chunk_sumcheck_chains_full_size(ns, ell):
    from nightstream_amd import twist
    from nightstream_amd.sidecar import chunk_lanes, chunk_rounds
    print(f"sidecar chunk of 2^{ell} cycles, batched driver", flush=True)
    lanes = chunk_lanes(40 + ell, ell)
    rounds = chunk_rounds(twist.TsOracle, lanes, ell, seed=7 + ell, batch=twist)
    assert len(rounds) == 30
    for label, claim, per, chals in rounds:
        assert per and len(per) == len(chals)
        running = kadd(per[0][0], per[0][1]) if label.endswith(("/read", "/write")) else claim
        for vals, r in zip(per, chals):
            assert kadd(vals[0], vals[1]) == running, label
            running = at(vals, r)
