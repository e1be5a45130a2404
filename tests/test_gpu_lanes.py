"""Lane-level Pi_RLC / Pi_DEC entry points (ns_rlc, ns_dec) reproduce a shard proof's main lane.

Step 1 of a Twist/Shout shard folds 13 ME instances (the step's MCS claim and the 12 children of
step 0, whose planes are the balanced split of step 0's digits).  Fed those instances, their
rotation vectors (the first columns of the recorded rotations) and the witnesses, ns_rlc must
return the recorded parent and ns_dec the recorded children, with every public check passing.
"""
import numpy as np
import pytest

from nightstream_amd import proof as P
from nightstream_amd.bus_workload import gen_bus_shard
from nightstream_amd.workload import Q

pytestmark = pytest.mark.gpu
SEED = bytes([0x4E, 0x53] + [0x07] * 30)
D = 54


def _me_equal(a, b):
    return (np.array_equal(np.asarray(a.c, np.uint64), np.asarray(b.c, np.uint64)) and a.m_in == b.m_in
            and np.array_equal(np.asarray(a.X, np.uint64), np.asarray(b.X, np.uint64)) and list(a.r) == list(b.r)
            and [list(r) for r in a.y] == [list(r) for r in b.y] and list(a.y_scalars) == list(b.y_scalars)
            and bytes(a.fold_digest) == bytes(b.fold_digest))


def test_lane_entry_points_match_the_proof(ns, oracle):
    print("lane-level RLC/DEC on a 2^5-cycle Twist/Shout shard", flush=True)
    inst, zs, mems, lay = gen_bus_shard(5, 2, 31)
    data, _ = oracle.Prover(inst, SEED).prove(zs, mem=mems)
    st = P.parse(data).steps[1]
    m = inst.m
    Z0 = oracle.decomp(zs[0])
    Z1 = oracle.decomp(zs[1])
    planes0, _ = oracle.split_b(Z0.astype(np.int32).reshape(-1), 12)
    wits = [Z1.reshape(-1)] + [np.asarray(p, np.int8).reshape(-1) for p in planes0]
    a = np.array([[v if v <= Q // 2 else v - Q for v in (int(x) for x in R[:, 0])] for R in st.rlc_rhos], np.int8)
    sess = ns.FoldingSession(ns.Ccs(inst), ns.PublicParams(m, seed=SEED))
    zb = [ns.DeviceBuffer.from_array(w) for w in wits]
    zmix = ns.DeviceBuffer(4 * D * m)
    parent = ns.rlc(sess, st.ccs_out, a, zb, zmix)
    assert _me_equal(parent, st.rlc_parent)
    planes = [ns.DeviceBuffer(D * m) for _ in range(12)]
    children, ok = ns.dec(sess, parent, zmix, 12, planes, bus_chunk=lay.chunk)
    assert ok == (True, True, True)
    for got, want in zip(children, st.dec_children):
        assert _me_equal(got, want)
