"""ns_shard_verify (host-only, no device): the transcript-replaying shard verifier of the library.

Proofs come from the CPU restatement (byte-identical to the MI355X prover's, as the GPU parity
tests show), so these tests run without a GPU.  The verifier must
  - accept honest CCS-only and Route A shard proofs (both obligation lanes);
  - reject every sampled single-word mutation, including each recorded challenge;
  - reject a forgery that the challenge-trusting algebraic checks accept: the last Ajtai round
    re-run at a challenge of the prover's choosing, with its round polynomial and the final sum
    solved so that every chain and the terminal identity still hold;
  - reject proofs bound to other public inputs, other memory metadata or other MCS commitments.
"""
import numpy as np
import pytest

from nightstream_amd import proof as proof_mod
from nightstream_amd import verify as algebra
from nightstream_amd.bus_workload import gen_bus_shard
from nightstream_amd.workload import Q, gen_r1cs_steps

SEED = bytes([0x4E, 0x53] + [0x07] * 30)


@pytest.fixture(scope="module")
def ns():
    import nightstream_amd
    return nightstream_amd


@pytest.fixture(scope="module")
def ccs_shard(oracle):
    inst, zs = gen_r1cs_steps(1 << 10, 9, 3)
    data, _ = oracle.Prover(inst, SEED).prove(zs)
    return inst, zs, data


@pytest.fixture(scope="module")
def mem_shard(oracle):
    inst, zs, mems, lay = gen_bus_shard(5, 2, 17)
    data, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems)
    return inst, zs, mems, data


def test_accepts_honest_proofs(ns, ccs_shard, mem_shard):
    inst, zs, data = ccs_shard
    ns.shard_verify(ns.Ccs(inst), zs, data)
    inst, zs, mems, data = mem_shard
    ns.shard_verify(ns.Ccs(inst), zs, data, mem=mems)


def test_rejects_single_word_mutations(ns, ccs_shard):
    inst, zs, data = ccs_shard
    ccs = ns.Ccs(inst)
    w = np.frombuffer(data, dtype="<u8").copy()
    rng = np.random.default_rng(7)
    for idx in rng.integers(1, len(w), size=40):
        w2 = w.copy()
        w2[idx] = (int(w2[idx]) + 1 + int(rng.integers(0, 1000))) % Q
        with pytest.raises(ns.NsError):
            ns.shard_verify(ccs, zs, w2.tobytes())


def test_rejects_each_recorded_challenge(ns, ccs_shard):
    """Every challenge the prover records (alpha, beta, gamma, each round challenge) is re-derived."""
    inst, zs, data = ccs_shard
    ccs = ns.Ccs(inst)
    pr = proof_mod.parse(data)
    w = np.frombuffer(data, dtype="<u8").copy()
    st = pr.steps[1]
    for ch in [st.alpha[2], st.beta_a[0], st.beta_r[3], st.gamma, st.sumcheck_challenges[4]]:
        pos = _find(w, [ch[0], ch[1]])
        w2 = w.copy()
        w2[pos] = (int(w2[pos]) + 1) % Q
        with pytest.raises(ns.NsError):
            ns.shard_verify(ccs, zs, w2.tobytes())


def _find(w, seq):
    seq = np.array(seq, dtype=np.uint64)
    hits = [i for i in np.nonzero(w == seq[0])[0] if i + len(seq) <= len(w) and np.array_equal(w[i:i + len(seq)], seq)]
    assert hits, "sequence not found"
    return int(hits[0])


def _kinv(a):
    # (a0 + a1 u)^-1 = (a0 - a1 u) / (a0^2 - 7 a1^2)
    n = (a[0] * a[0] - 7 * a[1] * a[1]) % Q
    ni = pow(n, Q - 2, Q)
    return (a[0] * ni % Q, (-a[1]) * ni % Q)


def test_rejects_prover_chosen_challenge_with_consistent_algebra(ns, ccs_shard):
    inst, zs, data = ccs_shard
    pr = proof_mod.parse(data)
    ell_n = 10
    s = 2
    st = pr.steps[s]
    acc = pr.steps[s - 1].dec_children
    rounds, chals = st.sumcheck_rounds, st.sumcheck_challenges
    run = st.sc_initial_sum
    for co, r in zip(rounds[:-1], chals[:-1]):
        run = algebra.poly_eval(co, r)
    last = [tuple(c) for c in rounds[-1]]
    r_new = ((chals[-1][0] + 1) % Q, chals[-1][1])
    alpha_p = chals[ell_n:-1] + [r_new]
    rhs = algebra._terminal_rhs(algebra.R1CS_F, st.ccs_out, chals[:ell_n], alpha_p, st.alpha, st.beta_a, st.beta_r,
                                st.gamma, acc[0].r)
    # solve c0, c1: 2 c0 + c1 = run - sum_{i>=2} c_i and c0 + c1 r = rhs - sum_{i>=2} c_i r^i
    tail01 = (0, 0)
    tailr = (0, 0)
    rp = algebra.K1
    for i, c in enumerate(last):
        if i >= 2:
            tail01 = algebra.kadd(tail01, c)
            tailr = algebra.kadd(tailr, algebra.kmul(c, rp))
        rp = algebra.kmul(rp, r_new)
    A = algebra.ksub(run, tail01)
    B = algebra.ksub(rhs, tailr)
    c0 = algebra.kmul(algebra.ksub(B, algebra.kmul(A, r_new)),
                      _kinv(algebra.ksub(algebra.K1, algebra.kscale(r_new, 2))))
    c1 = algebra.ksub(A, algebra.kscale(c0, 2))
    new_last = [c0, c1] + last[2:]
    assert algebra.kadd(algebra.poly_eval(new_last, algebra.K0), algebra.poly_eval(new_last, algebra.K1)) == run
    assert algebra.poly_eval(new_last, r_new) == rhs
    w = np.frombuffer(data, dtype="<u8").copy()
    pos = _find(w, [x for c in last for x in c])
    w[pos:pos + 4] = [c0[0], c0[1], c1[0], c1[1]]
    pos = _find(w, list(chals[-1]))
    w[pos:pos + 2] = list(r_new)
    pos = _find(w, list(st.sumcheck_final))
    w[pos:pos + 2] = list(rhs)
    forged = w.tobytes()
    # the challenge-trusting algebra accepts it ...
    assert algebra.verify_shard(proof_mod.parse(forged), algebra.R1CS_F, ell_n=ell_n)
    # ... the transcript replay does not
    with pytest.raises(ns.NsError, match="transcript"):
        ns.shard_verify(ns.Ccs(inst), zs, forged)


def test_rejects_other_statements(ns, mem_shard, ccs_shard):
    inst, zs, mems, data = mem_shard
    ccs = ns.Ccs(inst)
    xs = [z.copy() for z in zs]
    xs[0][0] = 2
    with pytest.raises(ns.NsError):
        ns.shard_verify(ccs, xs, data, mem=mems)
    # other memory metadata: one more initial RAM word
    import copy
    mems2 = copy.deepcopy(mems)
    mems2[1].mems[2].init = sorted(mems2[1].mems[2].init + [(65535, 5)]) if (65535, 5) not in mems2[1].mems[2].init \
        else mems2[1].mems[2].init[:-1]
    with pytest.raises(ns.NsError):
        ns.shard_verify(ccs, zs, data, mem=mems2)
    # a CCS-only reading of a memory proof
    with pytest.raises(ns.NsError):
        ns.shard_verify(ccs, zs, data)
    # the expected MCS commitments must match
    inst2, zs2, data2 = ccs_shard
    pr = proof_mod.parse(data2)
    mc = np.stack([st.ccs_out[0].c for st in pr.steps])
    ns.shard_verify(ns.Ccs(inst2), zs2, data2, mcs_commitments=mc)
    mc[1, 5] = (int(mc[1, 5]) + 1) % Q
    with pytest.raises(ns.NsError):
        ns.shard_verify(ns.Ccs(inst2), zs2, data2, mcs_commitments=mc)
