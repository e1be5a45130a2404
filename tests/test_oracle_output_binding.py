"""Output binding (fold_shard_prove_with_output_binding, neo-fold/src/shard.rs:2352-2385; the output
sum-check of neo-memory/src/output_check.rs) in the CPU restatement, checked by the library's
host-only verifier (ns_shard_verify_ob) and by what the protocol fixes:
  - honest proofs binding the RAM (num_bits 16) or the register file (num_bits 5) are accepted; they
    carry num_bits rounds of 4 coefficients, and the first round sums to zero;
  - the last step's batched time claims end with "output_binding/inc_total" of degree 2 + num_bits,
    and the verifier's closed forms hold at the output sum-check's point r': the inc_total claim's
    terminal value is has_write * inc * eq(wa_bits, r') of the Twist openings at r_time, and
    eq(r_addr, r') * io_mask(r') * (val_init(r') + inc_total - val_io(r')) is the output
    sum-check's final value (a wrong variable order, a dropped factor or a wrong sign fails them);
  - the verifier wants the binding exactly when the proof carries it;
  - a wrong claimed value is refused (the output sum-check no longer starts from zero);
  - a final state that disagrees with the trace at a claimed address, with claims that agree with
    it, is refused by the link to the Twist increments (val_init(r') + inc_total != val_final(r'));
  - the steps before the last are the same words with and without the binding;
  - invalid configurations (address width, address range, order, instance index) are refused.
"""
import numpy as np
import pytest

from nightstream_amd import proof as P
from nightstream_amd.bus_workload import gen_bus_shard
from nightstream_amd.workload import Q

SEED = bytes([0x4F, 0x42] + [0x05] * 30)
RAM, REG = 2, 0


@pytest.fixture(scope="module")
def ns():
    import nightstream_amd
    return nightstream_amd


@pytest.fixture(scope="module")
def shard():
    inst, zs, mems, lay, fin = gen_bus_shard(5, 2, 23, with_final=True)
    return inst, zs, mems, fin


def claims_at(final, addrs):
    return [(int(a), int(final[a])) for a in sorted(set(int(a) for a in addrs))]


def ram_claims(final, rng):
    written = np.nonzero(final)[0]
    picks = list(rng.choice(written, 6, replace=False)) + list(rng.integers(0, len(final), 3))
    return claims_at(final, picks)


def ob_dict(mem_idx, num_bits, claims, final):
    return {"mem_idx": mem_idx, "num_bits": num_bits, "claims": claims, "final_state": final}


def cfg(ns, mem_idx, num_bits, claims):
    io = ns.ProgramIO()
    for a, v in claims:
        io = io.with_output(a, v)
    return ns.OutputBindingConfig(num_bits, io, mem_idx)


@pytest.fixture(scope="module")
def ram_proof(oracle, shard):
    inst, zs, mems, fin = shard
    claims = ram_claims(fin[RAM], np.random.default_rng(3))
    data, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems, ob=ob_dict(RAM, 16, claims, fin[RAM]))
    return claims, data


def test_accepts_honest_bindings(ns, oracle, shard, ram_proof):
    inst, zs, mems, fin = shard
    claims, data = ram_proof
    ns.shard_verify(ns.Ccs(inst), zs, data, mem=mems, ob=cfg(ns, RAM, 16, claims))
    sp = P.parse(data)
    assert len(sp.output_rounds) == 16 and all(len(c) == 4 for c in sp.output_rounds)
    c0 = sp.output_rounds[0]
    assert ((2 * c0[0][0] + c0[1][0] + c0[2][0] + c0[3][0]) % Q, (2 * c0[0][1] + c0[1][1] + c0[2][1] + c0[3][1]) % Q) \
        == (0, 0)
    last = sp.steps[-1]
    assert last.batched_labels[-1] == b"output_binding/inc_total"
    assert last.batched_degree_bounds[-1] == 2 + 16
    # the register file: every address, including ones no step wrote
    reg_claims = claims_at(fin[REG], range(32))
    data2, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems, ob=ob_dict(REG, 5, reg_claims, fin[REG]))
    ns.shard_verify(ns.Ccs(inst), zs, data2, mem=mems, ob=cfg(ns, REG, 5, reg_claims))
    assert len(P.parse(data2).output_rounds) == 5
    # no claims at all: the output sum-check is identically zero
    data3, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems, ob=ob_dict(REG, 5, [], fin[REG]))
    ns.shard_verify(ns.Ccs(inst), zs, data3, mem=mems, ob=cfg(ns, REG, 5, []))
    assert all(c == [(0, 0)] * 4 for c in P.parse(data3).output_rounds)


def test_binding_presence_must_match(ns, oracle, shard, ram_proof):
    inst, zs, mems, fin = shard
    claims, data = ram_proof
    ccs = ns.Ccs(inst)
    with pytest.raises(ns.NsError, match="output binding"):
        ns.shard_verify(ccs, zs, data, mem=mems)
    plain, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems)
    with pytest.raises(ns.NsError, match="output binding"):
        ns.shard_verify(ccs, zs, plain, mem=mems, ob=cfg(ns, RAM, 16, claims))
    # the first step is the same words either way; the last step's transcript differs
    a, b = P.parse(data), P.parse(plain)
    assert a.steps[0].sumcheck_challenges == b.steps[0].sumcheck_challenges
    assert a.steps[-1].sumcheck_challenges != b.steps[-1].sumcheck_challenges
    assert b.output_rounds is None


def test_rejects_wrong_claim(ns, oracle, shard, ram_proof):
    inst, zs, mems, fin = shard
    claims, data = ram_proof
    bad = list(claims)
    bad[2] = (bad[2][0], (bad[2][1] + 1) % Q)
    with pytest.raises(ns.NsError, match="output"):
        ns.shard_verify(ns.Ccs(inst), zs, data, mem=mems, ob=cfg(ns, RAM, 16, bad))
    # the prover given the wrong claim: its sum-check starts from a nonzero sum
    data_bad, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems, ob=ob_dict(RAM, 16, bad, fin[RAM]))
    with pytest.raises(ns.NsError, match="output sum-check round"):
        ns.shard_verify(ns.Ccs(inst), zs, data_bad, mem=mems, ob=cfg(ns, RAM, 16, bad))


def test_rejects_final_state_that_disagrees_with_the_trace(ns, oracle, shard, ram_proof):
    inst, zs, mems, fin = shard
    claims, _ = ram_proof
    lie = fin[RAM].copy()
    a0 = claims[0][0]
    lie[a0] = (int(lie[a0]) + 5) % Q
    lie_claims = claims_at(lie, [a for a, _ in claims])
    data, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems, ob=ob_dict(RAM, 16, lie_claims, lie))
    sp = P.parse(data)
    c0 = sp.output_rounds[0]
    assert ((2 * c0[0][0] + c0[1][0] + c0[2][0] + c0[3][0]) % Q) == 0  # consistent with the lie
    with pytest.raises(ns.NsError, match="final check"):
        ns.shard_verify(ns.Ccs(inst), zs, data, mem=mems, ob=cfg(ns, RAM, 16, lie_claims))


def test_rejects_invalid_configurations(oracle, shard):
    inst, zs, mems, fin = shard
    pr = oracle.Prover(inst, SEED)
    with pytest.raises(RuntimeError, match="num_bits"):
        pr.prove(zs, mem=mems, ob=ob_dict(RAM, 15, [], fin[RAM][:1 << 15]))
    with pytest.raises(RuntimeError, match="mem_idx"):
        pr.prove(zs, mem=mems, ob=ob_dict(7, 5, [], fin[REG]))
    with pytest.raises(RuntimeError, match="num_bits"):
        pr.prove(zs, mem=mems, ob=ob_dict(REG, 16, [], fin[RAM]))
    with pytest.raises(ValueError):
        pr.prove(zs, mem=mems, ob=ob_dict(REG, 5, [], fin[RAM]))
    with pytest.raises(RuntimeError, match="below 2\\^num_bits"):
        pr.prove(zs, mem=mems, ob=ob_dict(REG, 5, [(40, 0)], fin[REG]))
