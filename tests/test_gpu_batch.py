"""Batched and served proving: every proof is the same bytes alone and in the mix.

Six shards of 2^10 / 2^11 rows with a distinct witness per step, two of them Twist/Shout shards
(Route A memory over the shared CPU bus), are proved one at a time in one prover context, then
through three concurrent contexts, then through the Poisson server; all three must agree, and a
CCS-only and a Twist/Shout shard from the batched path must equal the CPU restatement's bytes."""
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


@pytest.mark.parametrize("idx", [0, 1])
def test_batched_shard_matches_oracle(shards, oracle, idx):
    """Shard 0 (CCS-only) and shard 1 (Twist/Shout, 2^4-cycle chunks) from the batched path
    equal the CPU restatement's proofs of the same instances."""
    from nightstream_amd.batch import ShardFactory, prove_batched
    sh = shards[idx]
    assert (sh.mem is not None) == (idx == 1)
    inst, zs, mems = ShardFactory.instance(sh.n, sh.steps, sh.seed, sh.mem is not None)
    print(f"CPU restatement of shard {idx}", flush=True)
    data, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems)
    bat, _ = prove_batched(shards, 3)
    assert bat[idx] == data


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


def test_shared_round_launches(ns, monkeypatch):
    """Eight CCS-only shards proved by eight concurrent contexts with shared fused-round launches
    give the plain path's bytes, and the shared launcher carried every fused round (rounds
    4 .. ell_n - 1 of every step: from 2^9 rows round 2 runs from the digits and round 3 unfused
    on the tables folded after it), in no more launches than rounds."""
    from nightstream_amd import batch
    from nightstream_amd.batch import ShardFactory, prove_batched, prove_sequential
    monkeypatch.setattr(batch, "SHARE_ROUNDS", True)
    fac = ShardFactory(SEED)
    specs = [(1 << (10 + i % 3), 2 + i % 3, 40 + i) for i in range(8)]
    shards = [fac.make(n, steps, seed) for n, steps, seed in specs]
    print("plain path, 8 CCS-only shards", flush=True)
    seq, _ = prove_sequential(shards)
    l0, r0 = ns.shared_round_stats()
    print("eight contexts, shared rounds", flush=True)
    bat, _ = prove_batched(shards, 8)
    l1, r1 = ns.shared_round_stats()
    assert bat == seq
    want = sum(steps * (n.bit_length() - 1 - 4) for n, steps, _ in specs)
    assert r1 - r0 == want
    assert 0 < l1 - l0 <= r1 - r0
    print(f"shared rounds: {r1 - r0} rounds in {l1 - l0} launches", flush=True)
    assert not ns.shared_rounds_enabled()
