"""GPU parity of output binding (fold_shard_prove_with_output_binding, neo-fold/src/shard.rs:2352-2385).

The MI355X prover runs the output sum-check as a device product oracle and adds the
"output_binding/inc_total" Twist claim to the last step's batched time sum-check; every byte of
the proof must equal the CPU restatement's, from host and from device-resident witnesses, and the
host verifier must accept it with the same binding and refuse it with another claimed value.
"""
import numpy as np
import pytest

from nightstream_amd import proof as P
from nightstream_amd.bus_workload import gen_bus_shard
from nightstream_amd.workload import Q

pytestmark = pytest.mark.gpu
SEED = bytes([0x4F, 0x42] + [0x05] * 30)


def _cfg(ns, mem_idx, num_bits, claims):
    io = ns.ProgramIO()
    for a, v in claims:
        io = io.with_output(a, v)
    return ns.OutputBindingConfig(num_bits, io, mem_idx)


@pytest.mark.parametrize("mem_idx,num_bits,log_chunk,steps", [(2, 16, 5, 2), (0, 5, 6, 3)])
def test_output_binding_bit_identical(ns, oracle, mem_idx, num_bits, log_chunk, steps):
    print(f"output binding: mem {mem_idx} ({num_bits} bits), 2^{log_chunk}-cycle chunks, {steps} steps", flush=True)
    inst, zs, mems, _, fin = gen_bus_shard(log_chunk, steps, 31 + log_chunk, with_final=True)
    final = fin[mem_idx]
    rng = np.random.default_rng(num_bits)
    if num_bits == 16:
        written = np.nonzero(final)[0]
        addrs = sorted(set(int(a) for a in rng.choice(written, 8, replace=False)) | {0, 65535})
    else:
        addrs = list(range(32))
    claims = [(a, int(final[a])) for a in addrs]
    ob = {"mem_idx": mem_idx, "num_bits": num_bits, "claims": claims, "final_state": final}
    want, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems, ob=ob)
    cfg = _cfg(ns, mem_idx, num_bits, claims)
    sess = ns.FoldingSession(ns.Ccs(inst), ns.PublicParams(inst.m, seed=SEED))
    for z, mm in zip(zs, mems):
        sess.add_step_z(z, mem=mm)
    got = sess.fold_and_prove_with_output_binding(cfg, final)
    assert len(got) == len(want)
    assert got == want
    sess2 = ns.FoldingSession(ns.Ccs(inst), ns.PublicParams(inst.m, seed=SEED))
    bufs = [ns.DeviceBuffer.from_array(z) for z in zs]
    assert sess2.prove_device([b.data_ptr() for b in bufs], mem=mems, ob=cfg, final_state=final) == want
    assert len(P.parse(got).output_rounds) == num_bits
    ccs = ns.Ccs(inst)
    ns.shard_verify(ccs, zs, got, mem=mems, ob=cfg)
    bad = list(claims)
    bad[1] = (bad[1][0], (bad[1][1] + 3) % Q)
    with pytest.raises(ns.NsError):
        ns.shard_verify(ccs, zs, got, mem=mems, ob=_cfg(ns, mem_idx, num_bits, bad))


def test_output_binding_refuses_bad_config(ns):
    inst, zs, mems, _, fin = gen_bus_shard(5, 1, 41, with_final=True)
    sess = ns.FoldingSession(ns.Ccs(inst), ns.PublicParams(inst.m, seed=SEED))
    for z, mm in zip(zs, mems):
        sess.add_step_z(z, mem=mm)
    with pytest.raises(ns.NsError, match="num_bits"):
        sess.fold_and_prove_with_output_binding(_cfg(ns, 2, 15, []), fin[2][:1 << 15])
    with pytest.raises(ns.NsError, match="mem_idx"):
        sess.fold_and_prove_with_output_binding(_cfg(ns, 5, 5, []), fin[0])
    with pytest.raises(ns.NsError, match="address"):
        sess.fold_and_prove_with_output_binding(_cfg(ns, 0, 5, [(33, 1)]), fin[0])
    with pytest.raises(ns.NsError, match="2\\^num_bits"):
        sess.fold_and_prove_with_output_binding(_cfg(ns, 0, 5, []), fin[2])
