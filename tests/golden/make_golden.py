"""Writes tests/golden/golden.npz: small input/output vectors of the fold-step path, computed by
the CPU restatement (oracle/).  The reference holds no fixed vectors for this path, so these are
the build's own regression fixtures: the oracle must keep reproducing them (tests/test_golden.py,
CPU) and the MI355X library must reproduce them bit for bit (same file, marker gpu).

Run from the repository root:  python tests/golden/make_golden.py
Every input is either stored or regenerated from a stored seed; a SHA-256 of each regenerated
input is stored too, so a change in the generators shows up as a fixture failure.
"""
import hashlib
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, "tests"))

from oracle import pyoracle as oracle  # noqa: E402
from nightstream_amd.workload import Q, gen_memory_trace, gen_r1cs, rand_balanced_z, rand_field  # noqa: E402

D = 54
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden.npz")


def sha(*arrays) -> np.ndarray:
    h = hashlib.sha256()
    for a in arrays:
        h.update(np.ascontiguousarray(a).tobytes())
    return np.frombuffer(h.digest(), np.uint8).copy()


def rand_k(rng, n):
    return rand_field(rng, 2 * n, nonzero=False).reshape(n, 2)


def r1cs_arrays(inst, prefix):
    out = {}
    for name, mat in (("a", inst.a), ("b", inst.b), ("c", inst.c)):
        rp, cols, vals = mat
        out[f"{prefix}_{name}_ptr"] = np.asarray(rp, np.uint64)
        out[f"{prefix}_{name}_col"] = np.asarray(cols, np.uint32)
        out[f"{prefix}_{name}_val"] = np.asarray(vals, np.uint64)
    out[f"{prefix}_z"] = np.asarray(inst.z, np.uint64)
    out[f"{prefix}_nm"] = np.array([inst.n, inst.m], np.uint64)
    return out


def ccs_round_vectors(g):
    """k = 3 Pi_CCS row rounds on a 64-row CCS, then precompute_for_r at the folded point."""
    n, ell_n, ell_d, k = 64, 6, 6, 3
    inst, zs = gen_r1cs(n, 3, 1)
    rng = np.random.default_rng(17)
    Zs = [oracle.decomp(zs[0])] + [rng.choice(np.array([-1, 0, 1], np.int8), size=(D, n)) for _ in range(k - 1)]
    alpha, beta_a, beta_r = rand_k(rng, ell_d), rand_k(rng, ell_d), rand_k(rng, ell_n)
    gamma = rand_k(rng, 1)[0]
    r_in = rand_k(rng, ell_n)
    rs = rand_k(rng, ell_n)
    occs = oracle.Ccs(inst)
    oo = oracle.RowState(occs, Zs, alpha, beta_a, beta_r, gamma, ell_d, ell_n, r_in)
    xs = np.array([[x, 0] for x in range(7)], np.uint64)
    evals = []
    for rd in range(ell_n):
        evals.append(np.asarray(oo.evals(xs), np.uint64))
        oo.fold(rs[rd])
    y = oracle.precompute_y(occs, Zs, rs, ell_n)
    g.update(r1cs_arrays(inst, "ccs"))
    g.update({"ccs_Zs": np.stack(Zs), "ccs_alpha": alpha, "ccs_beta_a": beta_a, "ccs_beta_r": beta_r,
              "ccs_gamma": gamma, "ccs_r_in": r_in, "ccs_rs": rs, "ccs_evals": np.stack(evals),
              "ccs_y_eval": np.asarray(y, np.uint64)})


def main():
    oracle.build()
    g = {}
    # balanced base-2 decomposition, including the edges of the representable range
    rng = np.random.default_rng(1)
    z = rand_balanced_z(rng, 64)
    z[1:9] = [0, 1, Q - 1, (1 << 54) - 1, Q - ((1 << 54) - 1), 2, Q - 2, 3]
    g["dec_z"] = z
    g["dec_digits"] = oracle.decomp(z)
    # PP from a ChaCha8 seed, then commitments of two digit matrices
    seed = np.arange(32, dtype=np.uint8)
    m = 128
    Zc = rng.choice(np.array([-1, 0, 1], np.int8), size=(2, D, m))
    pp = oracle.ajtai_setup(seed.tobytes(), 16, m)
    g["commit_seed"] = seed
    g["commit_m"] = np.array([m], np.uint64)
    g["commit_Z"] = Zc
    g["commit_c"] = np.stack([oracle.ajtai_commit(pp, 16, m, Zc[w]) for w in range(2)])
    # Pi_CCS rounds and y
    ccs_round_vectors(g)
    # whole shard: n = 2^10, 3 fold steps (proof stored as its length and SHA-256)
    inst, zs = gen_r1cs(1 << 10, 11, 3)
    pseed = bytes([5] * 32)
    proof, _ = oracle.Prover(inst, pseed).prove(zs)
    a, b, c = inst.a, inst.b, inst.c
    g["shard_params"] = np.array([1 << 10, 11, 3], np.uint64)
    g["shard_pp_seed"] = np.frombuffer(pseed, np.uint8).copy()
    g["shard_inputs_sha"] = sha(*a, *b, *c, *zs)
    g["shard_proof_len"] = np.array([len(proof)], np.uint64)
    g["shard_proof_sha"] = sha(np.frombuffer(proof, np.uint8))
    # a Twist read-check time lane (ell_n = 6, 3 address bits) driven by stored challenges
    tr = gen_memory_trace(6, 3, 1.0, 0.4, seed=21)
    rng = np.random.default_rng(23)
    r_cycle = rand_k(rng, 6)
    r_addr = rand_k(rng, 3)
    init_at = rand_k(rng, 1)[0]
    ev, rounds_rs, final, claim = twist_vectors(oracle.TsOracle, tr, r_cycle, r_addr, init_at, rng)
    g["twist_r_cycle"], g["twist_r_addr"], g["twist_init_at"], g["twist_rs"] = r_cycle, r_addr, init_at, rounds_rs
    g["twist_claim"] = claim
    g["twist_evals"] = ev
    g["twist_final"] = final
    g["twist_trace_sha"] = trace_sha(tr)
    np.savez_compressed(OUT, **g)
    print(f"wrote {OUT}: {os.path.getsize(OUT)} bytes, {len(g)} arrays")


def kl(a):
    return [(int(x[0]), int(x[1])) for x in np.asarray(a).reshape(-1, 2)]


def twist_vectors(cls, tr, r_cycle, r_addr, init_at, rng=None, rs=None):
    """Drive a twist_read oracle of class `cls` through all rounds: values at 0..deg per round,
    the final value at rs[0] and the claim.  rs is drawn from rng when not given."""
    to = cls("twist_read", kl(r_cycle), tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["inc"],
             tr["wa_bits"], kl(r_addr), kl(init_at)[0])
    deg, rounds = to.degree_bound(), to.num_rounds()
    if rs is None:
        rs = rand_k(rng, rounds)
    ev = []
    for rd in range(rounds):
        ev.append(np.asarray(to.evals([(v, 0) for v in range(deg + 1)]), np.uint64))
        to.fold(kl(rs[rd])[0])
    final = np.asarray(to.evals(kl(rs[0])), np.uint64)
    return np.stack(ev), rs, final, np.array(to.claim, np.uint64)


def trace_sha(tr) -> np.ndarray:
    def canon(v):
        if isinstance(v, np.ndarray):
            return canon(v.tolist())
        if isinstance(v, (list, tuple)):
            return [canon(x) for x in v]
        return int(v)
    import json
    blob = json.dumps([(k, canon(tr[k])) for k in sorted(tr)]).encode()
    return np.frombuffer(hashlib.sha256(blob).digest(), np.uint8).copy()


if __name__ == "__main__":
    main()
