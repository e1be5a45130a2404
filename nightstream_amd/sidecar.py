"""The memory sidecar of a shard step: Twist/Shout round oracles over a RISC-V-like trace.

A chunk of 2^ell_n cycles carries five memory lanes (neo-memory `memory.rs:232-621`): two
register-file Twists (32 words, one read port each, writes on the first), the RAM Twist (2^16
words, 35% of cycles, Zipf(1.1) addresses, 40% writes) and three Shout lookup lanes into the
8-bit x 8-bit ADD / XOR / AND tables (2^16 entries each).  Each Twist lane gives seven oracles
(read and write checks, val_eval, total_inc, bitness, and the read and write address lanes),
each Shout lane three (value, index adapter, address lookup): 30 oracles per chunk.

The oracles of a chunk are grouped by their number of rounds and each group is driven through
all of its rounds with shared challenges, as `run_batched_sumcheck_prover`
(`neo-reductions/src/sumcheck.rs:350-456`) drives a batch: with the batched driver every round
of a group is one `batch_round` / `batch_fold` pair (every oracle's kernels enqueued before the
host waits once per oracle); the plain driver evaluates and folds one oracle at a time.  Both
produce the same values.

This module drives the round oracles alone, at their real shape, for the oracle-level
benchmark and tests: its challenges and random points come from a seeded generator, so its
values bind nothing.  Proofs of shards with memory take the Route A path of the shard prover
(`ns_prove_shard_mem`), whose memory lanes draw every challenge from the step transcript.
"""
from __future__ import annotations

import hashlib

import numpy as np

from .workload import Q, gen_lookup_trace, gen_memory_trace


def chunk_lanes(seed: int, ell_n: int = 12):
    """The memory lanes of one chunk of 2^ell_n cycles: ([regs_a, regs_b, ram], [add, xor, and])."""
    regs_a = gen_memory_trace(ell_n, 5, 1.0, 0.5, seed=seed * 8 + 1)
    regs_b = gen_memory_trace(ell_n, 5, 1.0, 0.0, seed=seed * 8 + 2)
    ram = gen_memory_trace(ell_n, 16, 0.35, 0.4, seed=seed * 8 + 3, zipf_s=1.1)
    return [regs_a, regs_b, ram], gen_lookup_trace(ell_n, seed=seed * 8 + 4)


def _rand_k(rng):
    return (int(rng.integers(0, 1 << 63)) * 2 % Q, int(rng.integers(0, 1 << 63)) * 2 % Q)


def chunk_oracles(cls, twists, shouts, rng, ell_n: int = 12):
    """Every Twist/Shout oracle of one chunk (time lanes, address lanes, lookups) as (label,
    constructor) pairs over class `cls` (the GPU TsOracle or the CPU restatement's)."""
    out = []
    for i, tr in enumerate(twists):
        ea = tr["ell_addr"]
        r_cycle = [_rand_k(rng) for _ in range(ell_n)]
        r_addr = [_rand_k(rng) for _ in range(ea)]
        init_at = _rand_k(rng)
        r_time = [_rand_k(rng) for _ in range(ell_n)]
        w = [_rand_k(rng) for _ in tr["ra_bits"]]
        out += [
            (f"twist{i}/read", lambda tr=tr, a=(r_cycle, r_addr, init_at): cls(
                "twist_read", a[0], tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["inc"], tr["wa_bits"],
                a[1], a[2])),
            (f"twist{i}/write", lambda tr=tr, a=(r_cycle, r_addr, init_at): cls(
                "twist_write", a[0], tr["has_write"], tr["wv"], tr["inc"], tr["wa_bits"], a[1], a[2])),
            (f"twist{i}/val_eval", lambda tr=tr, a=(r_addr, r_time): cls(
                "val_eval", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], a[0], a[1])),
            (f"twist{i}/total_inc", lambda tr=tr, a=r_addr: cls(
                "total_inc", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], a)),
            (f"twist{i}/bitness", lambda tr=tr, a=(r_cycle, w): cls("bitness", a[0], tr["ra_bits"], a[1])),
            (f"twist{i}/read_addr", lambda tr=tr, a=r_cycle: cls(
                "read_addr", tr["init"], a, tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["wa_bits"],
                tr["inc"])),
            (f"twist{i}/write_addr", lambda tr=tr, a=r_cycle: cls(
                "write_addr", tr["init"], a, tr["has_write"], tr["wv"], tr["wa_bits"], tr["inc"])),
        ]
    for j, ln in enumerate(shouts):
        r_cycle = [_rand_k(rng) for _ in range(ell_n)]
        r_addr = [_rand_k(rng) for _ in range(16)]
        out += [
            (f"shout{j}/value", lambda ln=ln, a=r_cycle: cls("shout_value", a, ln["has_lookup"], ln["val"])),
            (f"shout{j}/index_adapter", lambda ln=ln, a=(r_cycle, r_addr): cls(
                "index_adapter", a[0], ln["has_lookup"], ln["addr_bits"], a[1])),
            (f"shout{j}/addr_lookup", lambda ln=ln, a=r_cycle: cls(
                "addr_lookup", ln["addr_bits"], ln["has_lookup"], ln["table"], a)),
        ]
    return out


def drive_group(os, chals, batch=None):
    """All rounds of oracles with the same number of rounds and shared challenges (the round loop
    of run_batched_sumcheck_prover): with `batch` (the GPU module) every round of the group is one
    batch_round / batch_fold; without it each oracle is evaluated at 0..deg and folded in turn.
    Returns per oracle the list of per-round values."""
    vals = [[] for _ in os]
    for rd in range(os[0].num_rounds() if os else 0):
        if batch is not None:
            for i, y in enumerate(batch.batch_round(os)):
                vals[i].append(y)
            batch.batch_fold(os, chals[rd])
        else:
            for i, o in enumerate(os):
                vals[i].append(o.evals([(v, 0) for v in range(o.degree_bound() + 1)]))
                o.fold(chals[rd])
    return vals


def step_seed(tag: bytes, step: int) -> int:
    """A 64-bit seed for one step's sidecar randomness, bound to `tag` (the step's CCS proof)."""
    return int.from_bytes(hashlib.sha256(tag + step.to_bytes(4, "little")).digest()[:8], "little")


def chunk_rounds(cls, lanes, ell_n: int, seed: int, batch=None):
    """Build a chunk's 30 oracles over class `cls` (random points from `seed`) and drive every
    group of equal round count through its rounds with challenges from `seed` (grouped rounds
    through module `batch`, or one oracle at a time when it is None).  Returns per oracle
    (label, claim, [values at 0..deg of each round], [the challenge of each round]); the claim
    is the constructor's (zero for the Twist read and write checks, whose constructors return
    none)."""
    twists, shouts = lanes
    rng = np.random.default_rng(seed)
    labeled = [(lbl, mk()) for lbl, mk in chunk_oracles(cls, twists, shouts, rng, ell_n)]
    made = [o for _, o in labeled]
    groups = {}
    for i, o in enumerate(made):
        groups.setdefault(o.num_rounds(), []).append(i)
    out = [None] * len(made)
    for nr in sorted(groups):
        idxs = groups[nr]
        chals = [_rand_k(rng) for _ in range(nr)]
        for i, v in zip(idxs, drive_group([made[i] for i in idxs], chals, batch)):
            out[i] = (labeled[i][0], made[i].claim, v, chals)
    return out


def chunk_values(cls, lanes, ell_n: int, seed: int, batch=None) -> bytes:
    """`chunk_rounds` as bytes: each oracle's claim and every round's values (K as two u64
    words, little-endian) in oracle order."""
    words = [w for _, claim, per, _ in chunk_rounds(cls, lanes, ell_n, seed, batch)
             for k in [claim] + [k for rd in per for k in rd] for w in k]
    return np.asarray(words, dtype=np.uint64).tobytes()


def prove_chunk(lanes, ell_n: int, seed: int, batched: bool = True) -> bytes:
    """One step's memory sidecar on the GPU (`chunk_values` over the device oracles): `batched`
    picks the grouped driver (one batch_round / batch_fold per group round) or the plain one."""
    from . import twist
    return chunk_values(twist.TsOracle, lanes, ell_n, seed, twist if batched else None)
