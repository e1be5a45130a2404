"""Batched and served proving: every proof is the same bytes alone and in the mix.

Six shards of 2^10 / 2^11 rows, two of them with the Twist/Shout memory sidecar, are proved one
at a time in one prover context (sidecar oracles one at a time), then through three concurrent
contexts (grouped sidecar rounds), then through the Poisson server; all three must agree, and
the CCS proof and a sidecar chunk must equal the CPU restatement's."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SEED = bytes([0x4E, 0x53] + [0x07] * 30)


@pytest.fixture(scope="module")
def shards(ns):
    from nightstream_amd.batch import ShardFactory
    fac = ShardFactory(SEED)
    specs = [(1 << 10, 3, 1), (1 << 11, 2, 2), (1 << 10, 2, 3), (1 << 11, 3, 4), (1 << 10, 4, 5), (1 << 11, 1, 6)]
    return [fac.make(n, steps, seed, memory=seed in (2, 5)) for n, steps, seed in specs]


def test_batched_equals_plain(shards):
    from nightstream_amd.batch import prove_batched, prove_sequential
    print("plain path", flush=True)
    seq, _ = prove_sequential(shards)
    print("three contexts", flush=True)
    bat, _ = prove_batched(shards, 3)
    assert seq == bat
    assert len(set(seq)) == len(seq)


def test_warm_pool_reused(shards):
    """A warmed pool (contexts created and sized up front) serves two batched runs without
    creating more contexts, and its proofs equal the plain path's."""
    from nightstream_amd.batch import ProverPool, prove_batched, prove_sequential
    print("warmed pool of three contexts, two runs", flush=True)
    seq, _ = prove_sequential(shards)
    pool = ProverPool(3)
    pool.warm(shards)
    assert pool.created == 3
    for _ in range(2):
        bat, _ = prove_batched(shards, 3, pool)
        assert bat == seq
    assert pool.created == 3


def test_rebind_reuses_buffers(shards):
    import nightstream_amd as ns
    a, b = shards[1], shards[0]
    s = ns.FoldingSession(a.ccs, a.pp)
    pa = s.prove_device(a.z_ptrs())
    s.rebind(b.ccs, b.pp)
    pb = s.prove_device(b.z_ptrs())
    s.rebind(a.ccs, a.pp)
    assert s.prove_device(a.z_ptrs()) == pa
    assert ns.FoldingSession(b.ccs, b.pp).prove_device(b.z_ptrs()) == pb


def test_served_equals_alone(shards):
    from nightstream_amd.batch import ProverPool, serve
    print("serve", flush=True)
    res = serve(shards, rate=50.0, duration=0.3, workers=3, hbm_cap=1e12, seed=1)
    assert res, "no arrivals"
    alone = ProverPool(1)
    for r in res:
        assert r.proof == alone.prove(shards[r.job % len(shards)])
        assert r.finish >= r.start >= r.arrival - 1e-3


def test_batched_shard_matches_oracle(shards, oracle):
    from nightstream_amd.batch import gen_r1cs_fast, prove_batched
    sh = shards[0]
    inst, z = gen_r1cs_fast(sh.n, sh.seed)
    data, _ = oracle.Prover(inst, SEED).prove([z] * sh.steps)
    bat, _ = prove_batched(shards, 3)
    assert bat[0] == data


def test_sidecar_chunk_matches_oracle(shards, oracle):
    """A memory shard's proof is its CCS proof followed by one sidecar chunk per step; the first
    chunk from the batched path equals the CPU restatement's oracles driven one at a time with
    the same points and challenges."""
    import nightstream_amd as ns
    from oracle import pyoracle
    from nightstream_amd.batch import prove_batched
    from nightstream_amd.sidecar import chunk_values, step_seed
    sh = shards[1]
    assert sh.mem is not None and sh.mem_ell == 11
    bat, _ = prove_batched(shards, 3)
    ccs = ns.FoldingSession(sh.ccs, sh.pp).prove_device(sh.z_ptrs())
    assert bat[1].startswith(ccs)
    pyoracle.build()
    print("CPU restatement of one chunk", flush=True)
    want = chunk_values(pyoracle.TsOracle, sh.mem, sh.mem_ell, step_seed(ccs, 0), None)
    assert len(bat[1]) == len(ccs) + sh.steps * len(want)
    assert bat[1][len(ccs):len(ccs) + len(want)] == want


def test_process_prover_equals_plain(shards):
    """Two prover processes of two contexts each, over the same specs, give the plain path's
    proofs (the processes build their own shards from the specs)."""
    from nightstream_amd.batch import ProcessProver, prove_sequential
    specs = [(sh.n, sh.steps, sh.seed, sh.mem is not None) for sh in shards]
    print("two prover processes of two contexts", flush=True)
    pp = ProcessProver(SEED, specs, 2, 2, hw_queues=1)
    try:
        out, _ = pp.prove_all(list(range(len(specs))))
    finally:
        pp.close()
    assert pp.created == 4
    seq, _ = prove_sequential(shards)
    assert out == seq
