"""Golden vectors (tests/golden/golden.npz, written by tests/golden/make_golden.py from the CPU
restatement): the oracle must keep reproducing them, and the MI355X library must reproduce them
bit for bit through the C ABI.  Covered: balanced decomposition, PP setup + Ajtai commitments,
Pi_CCS row rounds and precompute_for_r (k = 3, with ME inputs), a whole 3-step shard proof
(length and SHA-256), and a Twist read-check time lane."""
import hashlib
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "golden"))

from make_golden import kl, sha, trace_sha, twist_vectors  # noqa: E402

from nightstream_amd.workload import R1CS, gen_memory_trace, gen_r1cs  # noqa: E402

G = np.load(os.path.join(HERE, "golden", "golden.npz"), allow_pickle=False)


def ccs_inst():
    mats = []
    for name in "abc":
        mats.append((G[f"ccs_{name}_ptr"], G[f"ccs_{name}_col"], G[f"ccs_{name}_val"]))
    n, m = (int(x) for x in G["ccs_nm"])
    return R1CS(n, m, mats[0], mats[1], mats[2], G["ccs_z"])


def ccs_args():
    return ([G["ccs_Zs"][i] for i in range(G["ccs_Zs"].shape[0])], G["ccs_alpha"], G["ccs_beta_a"], G["ccs_beta_r"],
            G["ccs_gamma"], 6, 6, G["ccs_r_in"])


def shard_inputs():
    n, seed, steps = (int(x) for x in G["shard_params"])
    inst, zs = gen_r1cs(n, seed, steps)
    a, b, c = inst.a, inst.b, inst.c
    assert np.array_equal(sha(*a, *b, *c, *zs), G["shard_inputs_sha"]), "R1CS generator changed"
    return inst, zs, bytes(G["shard_pp_seed"])


def twist_trace():
    tr = gen_memory_trace(6, 3, 1.0, 0.4, seed=21)
    assert np.array_equal(trace_sha(tr), G["twist_trace_sha"]), "memory-trace generator changed"
    return tr


def check_twist(cls):
    ev, _, final, claim = twist_vectors(cls, twist_trace(), G["twist_r_cycle"], G["twist_r_addr"], G["twist_init_at"],
                                        rs=G["twist_rs"])
    assert np.array_equal(claim, G["twist_claim"]) or not G["twist_claim"].any()
    assert np.array_equal(ev, G["twist_evals"])
    assert np.array_equal(final, G["twist_final"])


# ---------------- the oracle reproduces its golden vectors (CPU) ----------------
def test_oracle_decomp_and_commit(oracle):
    assert np.array_equal(oracle.decomp(G["dec_z"]), G["dec_digits"])
    m = int(G["commit_m"][0])
    pp = oracle.ajtai_setup(G["commit_seed"].tobytes(), 16, m)
    for w in range(2):
        assert np.array_equal(oracle.ajtai_commit(pp, 16, m, G["commit_Z"][w]), G["commit_c"][w])


def test_oracle_ccs_rounds(oracle):
    occs = oracle.Ccs(ccs_inst())
    oo = oracle.RowState(occs, *ccs_args())
    xs = np.array([[x, 0] for x in range(7)], np.uint64)
    for rd in range(6):
        assert np.array_equal(np.asarray(oo.evals(xs), np.uint64), G["ccs_evals"][rd]), f"round {rd}"
        oo.fold(G["ccs_rs"][rd])
    y = oracle.precompute_y(occs, ccs_args()[0], G["ccs_rs"], 6)
    assert np.array_equal(np.asarray(y, np.uint64), G["ccs_y_eval"])


def test_oracle_shard_proof(oracle):
    inst, zs, seed = shard_inputs()
    proof, _ = oracle.Prover(inst, seed).prove(zs)
    assert len(proof) == int(G["shard_proof_len"][0])
    assert hashlib.sha256(proof).digest() == G["shard_proof_sha"].tobytes()


def test_oracle_twist_lane(oracle):
    check_twist(oracle.TsOracle)


# ---------------- the MI355X library reproduces them (GPU) ----------------
@pytest.mark.gpu
def test_gpu_decomp_and_commit(ns):
    assert np.array_equal(ns.decomp_b(G["dec_z"]), G["dec_digits"])
    m = int(G["commit_m"][0])
    pp = ns.PublicParams(m, seed=G["commit_seed"].tobytes())
    got = pp.commit([G["commit_Z"][w] for w in range(2)])
    for w in range(2):
        assert np.array_equal(got[w], G["commit_c"][w]), f"witness {w}"


@pytest.mark.gpu
def test_gpu_ccs_rounds(ns):
    go = ns.CcsOracle(ns.Ccs(ccs_inst()), *ccs_args())
    xs = np.array([[x, 0] for x in range(7)], np.uint64)
    for rd in range(6):
        assert np.array_equal(np.asarray(go.evals_at(xs), np.uint64), G["ccs_evals"][rd]), f"round {rd}"
        go.fold(G["ccs_rs"][rd])
    assert np.array_equal(np.asarray(go.y_eval(G["ccs_rs"]), np.uint64), G["ccs_y_eval"])


@pytest.mark.gpu
def test_gpu_shard_proof(ns):
    inst, zs, seed = shard_inputs()
    print("golden shard n=2^10 steps=3", flush=True)
    sess = ns.FoldingSession(ns.Ccs(inst), ns.PublicParams(inst.m, seed=seed))
    for z in zs:
        sess.add_step_z(z)
    proof = sess.fold_and_prove()
    assert len(proof) == int(G["shard_proof_len"][0])
    assert hashlib.sha256(proof).digest() == G["shard_proof_sha"].tobytes()


@pytest.mark.gpu
def test_gpu_twist_lane(ns):
    from nightstream_amd import twist
    check_twist(twist.TsOracle)
