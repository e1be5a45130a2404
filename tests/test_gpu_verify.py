"""The GPU prover's proofs pass the library's transcript-replaying verifier (ns_shard_verify),
at the timed CCS-only shape (n = 2^16) and at the Twist/Shout shard's full chunk (n = 2^19)."""
import pytest

from nightstream_amd.bus_workload import gen_bus_shard
from nightstream_amd.workload import gen_r1cs_steps

pytestmark = pytest.mark.gpu
SEED = bytes([0x4E, 0x53] + [0x07] * 30)


def test_gpu_ccs_shard_2e16_verifies(ns):
    print("verify: 2^16 CCS-only shard, 3 steps", flush=True)
    inst, zs = gen_r1cs_steps(1 << 16, 3, 3)
    ccs = ns.Ccs(inst)
    sess = ns.FoldingSession(ccs, ns.PublicParams(inst.m, seed=SEED))
    bufs = [ns.DeviceBuffer.from_array(z) for z in zs]
    proof = sess.prove_device([b.data_ptr() for b in bufs])
    ns.shard_verify(ccs, zs, proof)


def test_gpu_route_a_full_chunk_verifies(ns):
    print("verify: 2^12-cycle Twist/Shout shard, 2 steps (n = 2^19)", flush=True)
    inst, zs, mems, _ = gen_bus_shard(12, 2, 8)
    ccs = ns.Ccs(inst)
    sess = ns.FoldingSession(ccs, ns.PublicParams(inst.m, seed=SEED))
    bufs = [ns.DeviceBuffer.from_array(z) for z in zs]
    proof = sess.prove_device([b.data_ptr() for b in bufs], mem=mems)
    ns.shard_verify(ccs, zs, proof, mem=mems)
