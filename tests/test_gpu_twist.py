"""GPU Twist/Shout oracles against the CPU restatement, through the C ABI, on synthetic
RISC-V-like lanes: register file (32 words), RAM (2^16 words, Zipf(1.1) addresses, 35% of
cycles, 40% writes) and the 8x8-bit ADD/XOR/AND lookup table.  Every round's evaluations (at the
degree+1 nodes and a random point of K), every claim and the final values must be identical."""
import numpy as np
import pytest

from kmath import rand_k

from nightstream_amd.workload import gen_lookup_trace, gen_memory_trace

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def twist(ns):
    from nightstream_amd import twist
    return twist


def drive(g, c, rng, label):
    assert g.num_rounds() == c.num_rounds(), label
    assert g.degree_bound() == c.degree_bound(), label
    assert g.claim == c.claim or c.claim == (0, 0), label
    rounds = c.num_rounds()
    for rd in range(rounds):
        deg = c.degree_bound()
        xs = [(v, 0) for v in range(deg + 1)] + [rand_k(rng)]
        yc = c.evals(xs)
        yg = g.evals(xs)
        assert yg == yc, f"{label}: round {rd}"
        r = rand_k(rng)
        g.fold(r)
        c.fold(r)
    assert g.num_rounds() == c.num_rounds() == 0
    x = [rand_k(rng)]
    assert g.evals(x) == c.evals(x), f"{label}: final value"


def both(twist, oracle, kind, *args):
    return twist.TsOracle(kind, *args), oracle.TsOracle(kind, *args)


# 2^14 cycles, every one an access: a support of 16384 entries, above the one-workgroup limit of
# 8192, so the first rounds take the multi-kernel fold (writes are kept rare: the oracle's
# val_pre costs pairs x increment terms)
@pytest.mark.parametrize("ell_addr,ell_n,p,wf,zipf", [(5, 10, 1.0, 0.34, 0.0), (16, 10, 0.35, 0.4, 1.1),
                                                     (5, 14, 1.0, 0.01, 0.0)])
def test_twist_time_lanes(twist, oracle, ell_addr, ell_n, p, wf, zipf):
    tr = gen_memory_trace(ell_n, ell_addr, p, wf, seed=7 + ell_addr, zipf_s=zipf)
    rng = np.random.default_rng(ell_addr)
    r_cycle = [rand_k(rng) for _ in range(ell_n)]
    r_addr = [rand_k(rng) for _ in range(ell_addr)]
    init_at = rand_k(rng)
    print(f"twist time lanes ell_addr={ell_addr} ell_n={ell_n}", flush=True)
    g, c = both(twist, oracle, "twist_read", r_cycle, tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"],
                tr["inc"], tr["wa_bits"], r_addr, init_at)
    drive(g, c, rng, "twist_read")
    g, c = both(twist, oracle, "twist_write", r_cycle, tr["has_write"], tr["wv"], tr["inc"], tr["wa_bits"], r_addr,
                init_at)
    drive(g, c, rng, "twist_write")
    r_time = [rand_k(rng) for _ in range(ell_n)]
    g, c = both(twist, oracle, "val_eval", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], r_addr, r_time)
    drive(g, c, rng, "val_eval")
    g, c = both(twist, oracle, "total_inc", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], r_addr)
    drive(g, c, rng, "total_inc")
    w = [rand_k(rng) for _ in tr["ra_bits"]]
    g, c = both(twist, oracle, "bitness", r_cycle, tr["ra_bits"], w)
    drive(g, c, rng, "bitness")


def poly_at(vals, x):
    """Value at x of the polynomial through (j, vals[j]), j = 0..d (Lagrange, exact in K)."""
    from kmath import kadd, kinv, kmul, ksub
    acc = (0, 0)
    for j in range(len(vals)):
        num, den = (1, 0), (1, 0)
        for m in range(len(vals)):
            if m != j:
                num = kmul(num, ksub(x, (m, 0)))
                den = kmul(den, ksub((j, 0), (m, 0)))
        acc = kadd(acc, kmul(vals[j], kmul(num, kinv(den))))
    return acc


def test_twist_sumcheck_chain_large(twist):
    """2^15 cycles with ~11k increment terms, above the one-workgroup limit of 8192 for the fold
    and the val_pre prefix sums (the oracle restatement is quadratic there): every round must
    satisfy p(0) + p(1) = previous p(r), and the value after the last round (a separate kernel,
    the full val_pre sum) must equal the last p(r) -- the verifier's checks."""
    from kmath import kadd
    tr = gen_memory_trace(15, 5, 1.0, 0.34, seed=12)
    rng = np.random.default_rng(90)
    r_cycle = [rand_k(rng) for _ in range(15)]
    r_addr = [rand_k(rng) for _ in range(5)]
    init_at = rand_k(rng)
    for kind, args in (("twist_read", (r_cycle, tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["inc"],
                                       tr["wa_bits"], r_addr, init_at)),
                       ("twist_write", (r_cycle, tr["has_write"], tr["wv"], tr["inc"], tr["wa_bits"], r_addr,
                                        init_at))):
        print(f"{kind} 2^15 sum-check chain", flush=True)
        g = twist.TsOracle(kind, *args)
        claim = None
        for rd in range(g.num_rounds()):
            vals = g.evals([(v, 0) for v in range(g.degree_bound() + 1)])
            if claim is not None:
                assert kadd(vals[0], vals[1]) == claim, f"{kind}: round {rd}"
            r = rand_k(rng)
            claim = poly_at(vals, r)
            g.fold(r)
        assert g.evals([rand_k(rng)])[0] == claim, f"{kind}: final value"


# (5, 14): 16384 events, above the one-workgroup sort limit of 8192, so the device sort and scan run
@pytest.mark.parametrize("ell_addr,ell_n", [(5, 12), (16, 12), (5, 14)])
def test_twist_address_lanes(twist, oracle, ell_addr, ell_n):
    tr = gen_memory_trace(ell_n, ell_addr, 1.0 if ell_addr == 5 else 0.35, 0.4, seed=3 + ell_addr + ell_n,
                          zipf_s=0.0 if ell_addr == 5 else 1.1)
    rng = np.random.default_rng(50 + ell_addr)
    r_cycle = [rand_k(rng) for _ in range(ell_n)]
    print(f"twist address lanes ell_addr={ell_addr} ell_n={ell_n}", flush=True)
    g, c = both(twist, oracle, "read_addr", tr["init"], r_cycle, tr["has_read"], tr["rv"], tr["ra_bits"],
                tr["has_write"], tr["wa_bits"], tr["inc"])
    drive(g, c, rng, "read_addr")
    g, c = both(twist, oracle, "write_addr", tr["init"], r_cycle, tr["has_write"], tr["wv"], tr["wa_bits"], tr["inc"])
    drive(g, c, rng, "write_addr")


def test_shout_lanes(twist, oracle):
    ell_n = 12
    lanes = gen_lookup_trace(ell_n, seed=11)
    rng = np.random.default_rng(60)
    r_cycle = [rand_k(rng) for _ in range(ell_n)]
    for i, ln in enumerate(lanes):
        print(f"shout lane {i}", flush=True)
        r_addr = [rand_k(rng) for _ in range(16)]
        g, c = both(twist, oracle, "shout_value", r_cycle, ln["has_lookup"], ln["val"])
        drive(g, c, rng, "shout_value")
        g, c = both(twist, oracle, "index_adapter", r_cycle, ln["has_lookup"], ln["addr_bits"], r_addr)
        drive(g, c, rng, "index_adapter")
        g, c = both(twist, oracle, "addr_lookup", ln["addr_bits"], ln["has_lookup"], ln["table"], r_cycle)
        drive(g, c, rng, "addr_lookup")


def test_batched_sumcheck(twist, oracle):
    """The batched driver (run_batched_sumcheck_prover's round loop, shared challenges): the time
    lanes of a RAM Twist and a Shout lane (12 rounds each), then the address lanes of the same
    Twist and the Shout table (16 rounds each), each group stepped with ns_ts_batch_round /
    ns_ts_batch_fold on the GPU and one oracle at a time on the CPU; every value must be
    identical, and so must the final values."""
    ell_n = 12
    tr = gen_memory_trace(ell_n, 16, 0.35, 0.4, seed=31, zipf_s=1.1)
    ln = gen_lookup_trace(ell_n, seed=32)[1]
    rng = np.random.default_rng(80)
    rk = lambda n: [rand_k(rng) for _ in range(n)]  # noqa: E731
    r_cycle, r_addr, init_at, r_time, w, r_addr_s = rk(ell_n), rk(16), rand_k(rng), rk(ell_n), rk(16), rk(16)
    time_group = [
        ("twist_read", r_cycle, tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["inc"], tr["wa_bits"],
         r_addr, init_at),
        ("twist_write", r_cycle, tr["has_write"], tr["wv"], tr["inc"], tr["wa_bits"], r_addr, init_at),
        ("val_eval", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], r_addr, r_time),
        ("total_inc", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], r_addr),
        ("bitness", r_cycle, tr["ra_bits"], w),
        ("shout_value", r_cycle, ln["has_lookup"], ln["val"]),
        ("index_adapter", r_cycle, ln["has_lookup"], ln["addr_bits"], r_addr_s),
    ]
    addr_group = [
        ("read_addr", tr["init"], r_cycle, tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["wa_bits"],
         tr["inc"]),
        ("write_addr", tr["init"], r_cycle, tr["has_write"], tr["wv"], tr["wa_bits"], tr["inc"]),
        ("addr_lookup", ln["addr_bits"], ln["has_lookup"], ln["table"], r_cycle),
    ]
    for name, group in (("time lanes", time_group), ("address lanes", addr_group)):
        print(f"batched {name}: {len(group)} oracles", flush=True)
        gs = [twist.TsOracle(*a) for a in group]
        cs = [oracle.TsOracle(*a) for a in group]
        rounds = cs[0].num_rounds()
        assert all(g.num_rounds() == c.num_rounds() == rounds for g, c in zip(gs, cs))
        for rd in range(rounds):
            got = twist.batch_round(gs)
            for i, c in enumerate(cs):
                want = c.evals([(v, 0) for v in range(c.degree_bound() + 1)])
                assert got[i] == want, f"{name}: {group[i][0]} round {rd}"
            r = rand_k(rng)
            twist.batch_fold(gs, r)
            for c in cs:
                c.fold(r)
        x = [rand_k(rng)]
        for i, (g, c) in enumerate(zip(gs, cs)):
            assert g.num_rounds() == 0
            assert g.evals(x) == c.evals(x), f"{name}: {group[i][0]} final value"
        with pytest.raises(twist.NsError):
            twist.batch_round(gs)  # no rounds left
    # oracles with different numbers of rounds cannot share a batch (sumcheck.rs:357-368)
    mixed = [twist.TsOracle(*time_group[5]), twist.TsOracle(*addr_group[2])]
    with pytest.raises(twist.NsError):
        twist.batch_round(mixed)


def test_concurrent_construction(twist):
    """Oracles built on 8 host threads at once (the constructors release the GIL) give the same
    round values, claims and final values as the same oracles built one after another."""
    from concurrent.futures import ThreadPoolExecutor
    ell_n = 12
    tr = gen_memory_trace(ell_n, 16, 0.35, 0.4, seed=41, zipf_s=1.1)
    lanes = gen_lookup_trace(ell_n, seed=42)
    rng = np.random.default_rng(95)
    rk = lambda n: [rand_k(rng) for _ in range(n)]  # noqa: E731
    r_cycle, r_addr, init_at, r_time = rk(ell_n), rk(16), rand_k(rng), rk(ell_n)
    specs = [("twist_read", r_cycle, tr["has_read"], tr["rv"], tr["ra_bits"], tr["has_write"], tr["inc"],
              tr["wa_bits"], r_addr, init_at),
             ("twist_write", r_cycle, tr["has_write"], tr["wv"], tr["inc"], tr["wa_bits"], r_addr, init_at),
             ("val_eval", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], r_addr, r_time),
             ("total_inc", ell_n, tr["wa_bits"], tr["has_write"], tr["inc"], r_addr)]
    for ln in lanes:
        specs += [("shout_value", r_cycle, ln["has_lookup"], ln["val"]),
                  ("index_adapter", r_cycle, ln["has_lookup"], ln["addr_bits"], rk(16))]
    specs = specs * 2
    chals = rk(ell_n)

    def run(oracles):
        out = [[o.claim] for o in oracles]
        for rd in range(ell_n):
            for i, y in enumerate(twist.batch_round(oracles)):
                out[i].append(y)
            twist.batch_fold(oracles, chals[rd])
        x = [rand_k(np.random.default_rng(7))]
        for i, o in enumerate(oracles):
            out[i].append(o.evals(x))
        return out

    print(f"concurrent construction: {len(specs)} oracles on 8 threads", flush=True)
    with ThreadPoolExecutor(max_workers=8) as pool:
        par = list(pool.map(lambda a: twist.TsOracle(*a), specs))
    seq = [twist.TsOracle(*a) for a in specs]
    assert run(par) == run(seq)


def test_product_oracle(twist, oracle):
    rng = np.random.default_rng(70)
    facs = [[rand_k(rng) for _ in range(1 << 10)] for _ in range(3)]
    g, c = both(twist, oracle, "product", facs, 3)
    drive(g, c, rng, "product")


@pytest.mark.parametrize("chunk,bus_cols,steps_len", [(4096, 40, 4096), (4096, 40, 3000), (64, 7, 64)])
def test_bus_helpers(twist, oracle, chunk, bus_cols, steps_len):
    """decode / build_time_sparse_from_bus_col / append_bus_openings_to_me_instance on a 2^12-cycle
    chunk with 40 bus columns (two Twist lanes of 5 address bits, one Twist of 16, one Shout of 16)."""
    from kmath import D
    rng = np.random.default_rng(chunk + bus_cols + steps_len)
    m_in = 1
    m = bus_cols * chunk + 1024
    Z = rng.integers(-1, 2, size=(D, m)).astype(np.int8)
    Z[:, 1024:1024 + chunk // 2] = 0  # a run of zero cells: empty entries in the time columns
    bus_base = m - bus_cols * chunk
    ell_n = (m_in + chunk - 1).bit_length()
    print(f"bus chunk={chunk} cols={bus_cols} steps={steps_len}", flush=True)
    z = twist.bus_decode(Z)
    assert np.array_equal(z, oracle.bus_decode(Z))
    cols = twist.bus_time_cols(Z, m_in, bus_base, chunk, bus_cols, steps_len, ell_n)
    for c in range(bus_cols):
        ci, cv = oracle.bus_time_col(z, m_in, bus_base, chunk, bus_cols, c, steps_len, 1 << ell_n)
        assert np.array_equal(cols[c][0], ci) and cols[c][1] == cv, f"column {c}"
    r = [rand_k(rng) for _ in range(ell_n)]
    yg, ysg = twist.bus_openings(Z, m_in, bus_base, chunk, bus_cols, r)
    yc, ysc = oracle.bus_openings(Z, m_in, bus_base, chunk, bus_cols, r)
    assert yg == yc and ysg == ysc
