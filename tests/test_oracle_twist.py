"""Twist/Shout oracles of the CPU restatement against brute-force sums of their dense polynomials.

For every oracle of neo-memory/src/twist_oracle.rs the test writes down the summand as a product
of multilinear extensions of the (dense) columns and fresh factors, each computed independently
from its definition (chi_r, the MLE of [u < t] by summing over Boolean points, memory values by
replaying writes), and checks, round by round with random challenges:
  * the claimed sum returned at construction equals the sum over the Boolean hypercube,
  * evals_at(x) equals sum_{rest in {0,1}} summand(r_0..r_{j-1}, x, rest) at deg+1 nodes and a
    random point of K,
  * after the last round the oracle's value equals summand(r).
This is the identity the reference tests (ts_oracle_identity_tests.rs,
track_a_sparse_equivalence.rs) rely on.
"""
import itertools

import numpy as np
import pytest

from kmath import K0, K1, Q, chi, eq_points, kadd, kmul, ksub, mle, rand_k

ELL_N = 4
ELL_A = 3
T = 1 << ELL_N
A = 1 << ELL_A


def keq(a, b):
    return eq_points([a], [b])


def dense(col):
    d = [K0] * T
    for i, v in zip(col[0], col[1]):
        d[int(i)] = v
    return d


def sparse_from_dense(d):
    idx = [i for i, v in enumerate(d) if v != K0]
    return (np.array(idx, np.uint64), [d[i] for i in idx])


def bits_cols(addrs, active):
    """Boolean bit columns: 1 at time t where `active[t]` and bit b of addrs[t] is set."""
    cols = []
    for b in range(ELL_A):
        cols.append(sparse_from_dense([(1, 0) if active[t] and (addrs[t] >> b) & 1 else K0 for t in range(T)]))
    return cols


def mk_trace(rng, gate_random=False):
    has_r = [bool(rng.integers(0, 2)) for _ in range(T)]
    has_w = [bool(rng.integers(0, 2)) for _ in range(T)]
    ra = [int(rng.integers(0, A)) for _ in range(T)]
    wa = [int(rng.integers(0, A)) for _ in range(T)]

    def g(flag):
        if not flag:
            return K0
        return rand_k(rng) if gate_random else K1

    return dict(
        has_read=[g(f) for f in has_r], has_write=[g(f) for f in has_w],
        rv=[rand_k(rng) if f else K0 for f in has_r], wv=[rand_k(rng) if f else K0 for f in has_w],
        inc=[rand_k(rng) if f else K0 for f in has_w], ra=ra, wa=wa, has_r=has_r, has_w=has_w)


def run_rounds(o, summand, nvars, rng, claim=None):
    """Drive the oracle through all rounds against the brute-force summand."""
    total = K0
    for pt in itertools.product([K0, K1], repeat=nvars):
        total = kadd(total, summand(list(pt)))
    if claim is not None:
        assert claim == total
    chals = []
    running = total
    assert o.num_rounds() == nvars
    for rd in range(nvars):
        deg = o.degree_bound()
        xs = [(v, 0) for v in range(deg + 1)] + [rand_k(rng)]
        ys = o.evals(xs)
        rest = nvars - rd - 1
        for x, y in zip(xs, ys):
            exp = K0
            for pt in itertools.product([K0, K1], repeat=rest):
                exp = kadd(exp, summand(chals + [x] + list(pt)))
            assert y == exp, f"round {rd} at {x}"
        assert kadd(ys[0], ys[1]) == running
        r = rand_k(rng)
        # value of the round polynomial at r by Lagrange interpolation through the nodes 0..deg
        running = K0
        for i in range(deg + 1):
            num, den = K1, 1
            for j in range(deg + 1):
                if j != i:
                    num = kmul(num, ksub(r, (j, 0)))
                    den = den * (i - j) % Q
            running = kadd(running, kmul(kmul(ys[i], num), (pow(den, Q - 2, Q), 0)))
        o.fold(r)
        chals.append(r)
    assert o.num_rounds() == 0
    final = o.evals([rand_k(rng)])[0]
    assert final == summand(chals) == running


def lt_bool(u, s):
    return 1 if u < s else 0


def lt_mle_t(t_point, s_int):
    """MLE in t of [t < s] at t_point, summing over Boolean t."""
    w = chi(t_point)
    acc = K0
    for u in range(T):
        if u < s_int:
            acc = kadd(acc, w[u])
    return acc


@pytest.mark.parametrize("seed", [1, 2])
def test_shout_value(oracle, seed):
    rng = np.random.default_rng(100 + seed)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    gate = [K1 if rng.integers(0, 2) else K0 for _ in range(T)]
    val = [rand_k(rng) if rng.integers(0, 3) else K0 for _ in range(T)]
    o = oracle.TsOracle("shout_value", r_cycle, sparse_from_dense(gate), sparse_from_dense(val))
    f = lambda t: kmul(kmul(eq_points(t, r_cycle), mle(gate, t)), mle(val, t))
    run_rounds(o, f, ELL_N, rng, o.claim)


@pytest.mark.parametrize("gate_random", [False, True])
def test_index_adapter(oracle, gate_random):
    rng = np.random.default_rng(110 + gate_random)
    tr = mk_trace(rng, gate_random)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    r_addr = [rand_k(rng) for _ in range(ELL_A)]
    bits = bits_cols(tr["ra"], tr["has_r"])
    o = oracle.TsOracle("index_adapter", r_cycle, sparse_from_dense(tr["has_read"]), bits, r_addr)
    dbits = [dense(c) for c in bits]

    def f(t):
        acc = kmul(eq_points(t, r_cycle), mle(tr["has_read"], t))
        for b in range(ELL_A):
            acc = kmul(acc, keq(mle(dbits[b], t), r_addr[b]))
        return acc

    run_rounds(o, f, ELL_N, rng, o.claim)


def test_bitness(oracle):
    rng = np.random.default_rng(120)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    cols = []
    for c in range(3):
        # mostly Boolean, with a couple of non-bit values so the check is not trivially zero
        d = [K1 if rng.integers(0, 2) else K0 for _ in range(T)]
        d[int(rng.integers(0, T))] = rand_k(rng)
        cols.append(d)
    w = [rand_k(rng) for _ in cols]
    o = oracle.TsOracle("bitness", r_cycle, [sparse_from_dense(d) for d in cols], w)

    def f(t):
        s = K0
        for d, wi in zip(cols, w):
            b = mle(d, t)
            s = kadd(s, kmul(wi, kmul(b, ksub(b, K1))))
        return kmul(eq_points(t, r_cycle), s)

    run_rounds(o, f, ELL_N, rng)


def _inc_terms(tr, r_addr):
    ca = chi(r_addr)
    return {u: kmul(kmul(tr["has_write"][u], tr["inc"][u]), ca[tr["wa"][u]]) for u in range(T) if tr["has_w"][u]}


def _val_pre_dense(tr, r_addr, init_at):
    terms = _inc_terms(tr, r_addr)
    out = []
    for s in range(T):
        v = init_at
        for u, tv in terms.items():
            if u < s:
                v = kadd(v, tv)
        out.append(v)
    return out


@pytest.mark.parametrize("seed", [1, 2])
def test_twist_read_check(oracle, seed):
    rng = np.random.default_rng(130 + seed)
    tr = mk_trace(rng)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    r_addr = [rand_k(rng) for _ in range(ELL_A)]
    init_at = rand_k(rng)
    rbits = bits_cols(tr["ra"], tr["has_r"])
    wbits = bits_cols(tr["wa"], tr["has_w"])
    o = oracle.TsOracle("twist_read", r_cycle, sparse_from_dense(tr["has_read"]), sparse_from_dense(tr["rv"]), rbits,
                        sparse_from_dense(tr["has_write"]), sparse_from_dense(tr["inc"]), wbits, r_addr, init_at)
    vpd = _val_pre_dense(tr, r_addr, init_at)
    drb = [dense(c) for c in rbits]

    def f(t):
        acc = kmul(eq_points(t, r_cycle), mle(tr["has_read"], t))
        acc = kmul(acc, ksub(mle(vpd, t), mle(tr["rv"], t)))
        for b in range(ELL_A):
            acc = kmul(acc, keq(mle(drb[b], t), r_addr[b]))
        return acc

    run_rounds(o, f, ELL_N, rng)


def test_twist_write_check(oracle):
    rng = np.random.default_rng(140)
    tr = mk_trace(rng)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    r_addr = [rand_k(rng) for _ in range(ELL_A)]
    init_at = rand_k(rng)
    wbits = bits_cols(tr["wa"], tr["has_w"])
    o = oracle.TsOracle("twist_write", r_cycle, sparse_from_dense(tr["has_write"]), sparse_from_dense(tr["wv"]),
                        sparse_from_dense(tr["inc"]), wbits, r_addr, init_at)
    vpd = _val_pre_dense(tr, r_addr, init_at)
    dwb = [dense(c) for c in wbits]

    def f(t):
        acc = kmul(eq_points(t, r_cycle), mle(tr["has_write"], t))
        acc = kmul(acc, ksub(ksub(mle(tr["wv"], t), mle(vpd, t)), mle(tr["inc"], t)))
        for b in range(ELL_A):
            acc = kmul(acc, keq(mle(dwb[b], t), r_addr[b]))
        return acc

    run_rounds(o, f, ELL_N, rng)


def test_val_eval_and_total_inc(oracle):
    rng = np.random.default_rng(150)
    tr = mk_trace(rng)
    r_addr = [rand_k(rng) for _ in range(ELL_A)]
    r_time = [rand_k(rng) for _ in range(ELL_N)]
    wbits = bits_cols(tr["wa"], tr["has_w"])
    dwb = [dense(c) for c in wbits]
    # MLE in t of LT(t, r_time) = sum_{u, s} chi_t(u) chi_r(s) [u < s]
    cr = chi(r_time)

    def lt_t(t):
        ct = chi(t)
        acc = K0
        for u in range(T):
            for s in range(u + 1, T):
                acc = kadd(acc, kmul(ct[u], cr[s]))
        return acc

    def eqw(t):
        acc = K1
        for b in range(ELL_A):
            acc = kmul(acc, keq(mle(dwb[b], t), r_addr[b]))
        return acc

    o = oracle.TsOracle("val_eval", ELL_N, wbits, sparse_from_dense(tr["has_write"]), sparse_from_dense(tr["inc"]),
                        r_addr, r_time)
    f = lambda t: kmul(kmul(kmul(mle(tr["has_write"], t), mle(tr["inc"], t)), lt_t(t)), eqw(t))
    run_rounds(o, f, ELL_N, rng, o.claim)

    o2 = oracle.TsOracle("total_inc", ELL_N, wbits, sparse_from_dense(tr["has_write"]), sparse_from_dense(tr["inc"]),
                         r_addr)
    g = lambda t: kmul(kmul(mle(tr["has_write"], t), mle(tr["inc"], t)), eqw(t))
    run_rounds(o2, g, ELL_N, rng, o2.claim)


def _mem_before(tr, init, t):
    """Memory contents (dense over addresses) just before time t, replaying writes in order."""
    mem = [K0] * A
    for a, v in init:
        mem[a] = kadd(mem[a], v)
    for u in range(t):
        if tr["has_w"][u]:
            mem[tr["wa"][u]] = kadd(mem[tr["wa"][u]], kmul(tr["inc"][u], tr["has_write"][u]))
    return mem


def test_read_check_addr_lane(oracle):
    rng = np.random.default_rng(160)
    tr = mk_trace(rng)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    init = [(a, rand_k(rng)) for a in sorted(rng.choice(A, size=3, replace=False).tolist())]
    rbits = bits_cols(tr["ra"], tr["has_r"])
    wbits = bits_cols(tr["wa"], tr["has_w"])
    o = oracle.TsOracle("read_addr", init, r_cycle, sparse_from_dense(tr["has_read"]), sparse_from_dense(tr["rv"]),
                        rbits, sparse_from_dense(tr["has_write"]), wbits, sparse_from_dense(tr["inc"]))
    ct = chi(r_cycle)
    mems = [_mem_before(tr, init, t) for t in range(T)]

    def f(a_pt):
        ca = chi(a_pt)
        acc = K0
        for t in range(T):
            if not tr["has_r"][t]:
                continue
            ra_bits = [(1, 0) if (tr["ra"][t] >> b) & 1 else K0 for b in range(ELL_A)]
            val = K0
            for a in range(A):
                val = kadd(val, kmul(ca[a], mems[t][a]))
            term = kmul(kmul(ct[t], tr["has_read"][t]), eq_points(a_pt, ra_bits))
            acc = kadd(acc, kmul(term, ksub(val, tr["rv"][t])))
        return acc

    run_rounds(o, f, ELL_A, rng)


def test_write_check_addr_lane(oracle):
    rng = np.random.default_rng(170)
    tr = mk_trace(rng)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    init = [(a, rand_k(rng)) for a in sorted(rng.choice(A, size=2, replace=False).tolist())]
    wbits = bits_cols(tr["wa"], tr["has_w"])
    o = oracle.TsOracle("write_addr", init, r_cycle, sparse_from_dense(tr["has_write"]), sparse_from_dense(tr["wv"]),
                        wbits, sparse_from_dense(tr["inc"]))
    ct = chi(r_cycle)
    mems = [_mem_before(tr, init, t) for t in range(T)]

    def f(a_pt):
        ca = chi(a_pt)
        acc = K0
        for t in range(T):
            if not tr["has_w"][t]:
                continue
            wa_bits = [(1, 0) if (tr["wa"][t] >> b) & 1 else K0 for b in range(ELL_A)]
            val = K0
            for a in range(A):
                val = kadd(val, kmul(ca[a], mems[t][a]))
            term = kmul(kmul(ct[t], tr["has_write"][t]), eq_points(a_pt, wa_bits))
            acc = kadd(acc, kmul(term, ksub(ksub(tr["wv"][t], val), tr["inc"][t])))
        return acc

    run_rounds(o, f, ELL_A, rng)


def test_address_lookup_and_product(oracle):
    rng = np.random.default_rng(180)
    tr = mk_trace(rng)
    r_cycle = [rand_k(rng) for _ in range(ELL_N)]
    table = [rand_k(rng) for _ in range(A - 2)]  # shorter than 2^ell_addr: zero padded
    bits = bits_cols(tr["ra"], tr["has_r"])
    o = oracle.TsOracle("addr_lookup", bits, sparse_from_dense(tr["has_read"]), table, r_cycle)
    ct = chi(r_cycle)
    W = [K0] * A
    for t in range(T):
        if tr["has_r"][t]:
            W[tr["ra"][t]] = kadd(W[tr["ra"][t]], kmul(ct[t], tr["has_read"][t]))
    tab = table + [K0] * (A - len(table))
    f = lambda a_pt: kmul(mle(tab, a_pt), mle(W, a_pt))
    run_rounds(o, f, ELL_A, rng, o.claim)

    facs = [[rand_k(rng) for _ in range(8)] for _ in range(3)]
    o2 = oracle.TsOracle("product", facs, 3)
    g = lambda pt: kmul(kmul(mle(facs[0], pt), mle(facs[1], pt)), mle(facs[2], pt))
    run_rounds(o2, g, 3, rng)
