"""GPU parity of shards with Twist/Shout memory over the shared CPU bus (Route A).

Each step carries three Shout lookup lanes, two register-file Twists and the RAM Twist
(nightstream_amd/bus_workload.py).  The proof holds the address-lane sum-checks, the memory time
lanes of the batched time sum-check, the val-eval batch, the CPU ME claims at r_val, the main and
val RLC/DEC lanes and both obligation lanes; every byte must equal the CPU restatement's.
"""
import numpy as np
import pytest

from nightstream_amd import proof as P
from nightstream_amd.bus_workload import gen_bus_shard
from nightstream_amd.workload import Q

pytestmark = pytest.mark.gpu
SEED = bytes([0x4E, 0x53] + [0x07] * 30)


def _prove_gpu(ns, inst, zs, mems, device: bool):
    sess = ns.FoldingSession(ns.Ccs(inst), ns.PublicParams(inst.m, seed=SEED))
    if device:
        bufs = [ns.DeviceBuffer.from_array(z) for z in zs]
        return sess.prove_device([b.data_ptr() for b in bufs], mem=mems)
    for z, mm in zip(zs, mems):
        sess.add_step_z(z, mem=mm)
    return sess.fold_and_prove()


@pytest.mark.parametrize("log_chunk,steps", [(5, 3), (6, 2)])
def test_route_a_shard_bit_identical(ns, oracle, log_chunk, steps):
    print(f"Route A shard: 2^{log_chunk}-cycle chunks, {steps} steps", flush=True)
    inst, zs, mems, _ = gen_bus_shard(log_chunk, steps, 11 + log_chunk)
    want, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems)
    got = _prove_gpu(ns, inst, zs, mems, device=False)
    assert len(got) == len(want)
    assert got == want
    assert _prove_gpu(ns, inst, zs, mems, device=True) == want
    sp = P.parse(got)
    assert len(sp.final_main) == 12 and len(sp.final_val) == 12 * steps


def _chain_ok(rounds, chals, claimed):
    """p(0) + p(1) equals the claimed sum, then each previous p(r); returns the final value."""
    run = claimed
    for co, r in zip(rounds, chals):
        s01 = (2 * co[0][0] + sum(c[0] for c in co[1:])) % Q, (2 * co[0][1] + sum(c[1] for c in co[1:])) % Q
        assert s01 == tuple(run)
        acc = (0, 0)
        for c in reversed(co):
            acc = ((acc[0] * r[0] + 7 * acc[1] * r[1] + c[0]) % Q, (acc[0] * r[1] + acc[1] * r[0] + c[1]) % Q)
        run = acc
    return run


def test_route_a_full_chunk_sumcheck_chains(ns):
    """The config's 2^12-cycle chunk (n = m = 2^19): two steps on the GPU (the second with the Twist
    rollover claims); every batched time claim, address-lane claim and val-eval claim forms a valid
    sum-check chain under the proof's own challenges, and both obligation lanes are present."""
    print("Route A shard: 2^12-cycle chunks, 2 steps (n = 2^19)", flush=True)
    inst, zs, mems, _ = gen_bus_shard(12, 2, 5)
    data = _prove_gpu(ns, inst, zs, mems, device=True)
    sp = P.parse(data)
    for st in sp.steps:
        r_time = st.sumcheck_challenges[:19]
        assert st.batched_labels[0] == b"ccs/time" and len(st.batched_labels) == 19
        for rounds, claimed in zip(st.batched_round_polys, st.batched_claimed_sums):
            _chain_ok(rounds, r_time, claimed)
        for tw in st.mem.twist:
            for c in range(2):
                _chain_ok(tw.round_polys[c], tw.r_addr, (0, 0))
    assert len(sp.final_main) == 12 and len(sp.final_val) == 24


def test_route_a_rejects_unpadded_ccs(ns):
    """A CCS missing a padding row (1 - flag) * field = 0 of the bus is refused (the reference's
    shared-bus guardrail, memory_sidecar/cpu_bus.rs:723-939)."""
    inst, zs, mems, lay = gen_bus_shard(5, 1, 3)
    c = lay.chunk
    row = c  # the first padding row: (1 - has_lookup) * val of Shout lane 0 at j = 0
    a_ptr, a_col, a_val = inst.a
    lo, hi = int(a_ptr[row]), int(a_ptr[row + 1])
    a_col = np.delete(a_col, range(lo, hi))
    a_val = np.delete(a_val, range(lo, hi))
    a_ptr = a_ptr.copy()
    a_ptr[row + 1:] -= np.uint64(hi - lo)
    inst.a = (a_ptr, a_col, a_val)
    with pytest.raises(ns.NsError):
        _prove_gpu(ns, inst, zs, mems, device=False)
