"""The CPU restatement of the Route A memory sidecar (oracle/route_a.c) on small shared-bus shards.

Pins beyond the per-oracle brute-force checks of tests/test_oracle_twist.py:
  - every claim of every batched sum-check forms a valid chain under the proof's own challenges
    (checked here from the proof bytes, independently of the prover's in-loop checks);
  - the Twist read/write time claims equal the final values of the address-lane sum-checks, and
    the address lanes start from claimed sum zero;
  - a trace whose read value disagrees with memory (CCS still satisfied) is refused by the Twist
    read argument, and a lookup value that disagrees with the table is refused;
  - the val lane folds the current and previous steps' CPU witnesses (two claims from step 1 on)
    and both obligation lanes come out with 12 children per fold.
"""
import numpy as np
import pytest

from nightstream_amd import proof as P
from nightstream_amd.bus_workload import gen_bus_shard, r1cs_satisfied
from nightstream_amd.workload import Q

SEED = bytes([1] * 32)


def kev(co, r):
    acc = (0, 0)
    for c in reversed(co):
        acc = ((acc[0] * r[0] + 7 * acc[1] * r[1] + c[0]) % Q, (acc[0] * r[1] + acc[1] * r[0] + c[1]) % Q)
    return acc


def chain(rounds, chals, claimed):
    run = tuple(claimed)
    for co, r in zip(rounds, chals):
        s01 = ((2 * co[0][0] + sum(c[0] for c in co[1:])) % Q, (2 * co[0][1] + sum(c[1] for c in co[1:])) % Q)
        assert s01 == run
        run = kev(co, r)
    return run


@pytest.fixture(scope="module")
def shard(oracle):
    inst, zs, mems, lay = gen_bus_shard(5, 2, 21)
    assert all(r1cs_satisfied(inst, z) for z in zs)
    proof, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems)
    return inst, zs, mems, lay, P.parse(proof)


def test_route_a_chains_and_claims(shard):
    inst, zs, mems, lay, sp = shard
    ell_n = int(inst.n).bit_length() - 1
    for idx, st in enumerate(sp.steps):
        r_time = st.sumcheck_challenges[:ell_n]
        labels = [lb.decode() for lb in st.batched_labels]
        assert labels[0] == "ccs/time"
        assert labels[1:10] == ["shout/value", "shout/adapter", "shout/bitness"] * 3
        assert labels[10:] == ["twist/read_check", "twist/write_check", "twist/bitness"] * 3
        for rounds, claimed, deg in zip(st.batched_round_polys, st.batched_claimed_sums, st.batched_degree_bounds):
            assert all(len(co) == deg + 1 for co in rounds)
            chain(rounds, r_time, claimed)
        # bitness claims are zero; the Twist time claims are the address lanes' final values
        for i in (3, 6, 9, 12, 15, 18):
            assert tuple(st.batched_claimed_sums[i]) == (0, 0)
        for t, tw in enumerate(st.mem.twist):
            assert [tuple(x) for x in tw.claimed_sums] == [(0, 0), (0, 0)]
            fin_r = chain(tw.round_polys[0], tw.r_addr, (0, 0))
            fin_w = chain(tw.round_polys[1], tw.r_addr, (0, 0))
            assert tuple(st.batched_claimed_sums[10 + 3 * t]) == fin_r
            assert tuple(st.batched_claimed_sums[11 + 3 * t]) == fin_w
            ve = tw.val_eval
            assert (ve["prev_total"] is None) == (idx == 0)
        # the Shout address lanes: all three active, one shared r_addr of 16 coordinates
        assert st.mem.shout_active_mask == 0b111 and len(st.mem.shout_r_addr) == 16
        for rounds, claimed in zip(st.mem.shout_round_polys, st.mem.shout_claimed_sums):
            chain(rounds, st.mem.shout_r_addr, claimed)
        assert len(st.mem.cpu_me_claims_val) == (1 if idx == 0 else 2)
        assert st.val_fold is not None and len(st.val_fold[2]) == 12
    assert len(sp.final_main) == 12 and len(sp.final_val) == 24
    # every ME carries the core rows and one opening per bus column
    assert all(len(e.y) == 4 + lay.bus_cols for e in sp.final_main + sp.final_val)


def test_route_a_obligation_recompositions(shard):
    """DEC recomposition on both lanes, bus rows included: sum_i 2^i child_i = parent for every y
    row (the children's bus openings are openings of their planes, which recompose to Z_mix)."""
    inst, zs, mems, lay, sp = shard
    for st in sp.steps:
        for (_, parent, children) in [(st.rlc_rhos, st.rlc_parent, st.dec_children), st.val_fold]:
            for j in range(len(parent.y)):
                for rho in range(54):
                    acc = [0, 0]
                    for i, ch in enumerate(children):
                        acc[0] = (acc[0] + (ch.y[j][rho][0] << i)) % Q
                        acc[1] = (acc[1] + (ch.y[j][rho][1] << i)) % Q
                    assert tuple(acc) == tuple(parent.y[j][rho])


def test_route_a_rejects_inconsistent_memory(oracle):
    inst, zs, mems, lay = gen_bus_shard(5, 2, 22)
    ta = lay.twist[0]
    z = zs[1].copy()
    cell = lay.cell(ta[6], 9)  # rv of the rs1 register Twist at cycle 9
    z[cell] = (int(z[cell]) + 1) % Q
    assert r1cs_satisfied(inst, z)  # the CCS does not see it; the Twist read argument must
    with pytest.raises(RuntimeError, match="invariant"):
        oracle.Prover(inst, SEED).prove([zs[0], z], mem=mems)
    # a lookup result that is not the table's entry is refused
    z = zs[0].copy()
    (a0, e, hl, val) = lay.shout[0]
    j = int(np.nonzero(z[lay.cell(hl, np.arange(lay.chunk))])[0][0])
    z[lay.cell(val, j)] = (int(z[lay.cell(val, j)]) + 1) % Q
    with pytest.raises(RuntimeError):
        oracle.Prover(inst, SEED).prove([z], mem=[mems[0]])
