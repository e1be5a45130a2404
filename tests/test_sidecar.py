"""The memory-sidecar driver on the CPU restatement: every oracle of a small chunk runs a valid
sum-check chain under the driver's grouping and challenge order.

For each oracle, p_0(0) + p_0(1) equals its claim (where its constructor returns one), and
p_i(0) + p_i(1) equals p_{i-1} at the challenge of round i-1.  The round polynomials are given
by their values at 0..deg, so p(r) is the Lagrange interpolation over those integer nodes, in
K = F_q[u]/(u^2 - 7) with Python integers (neo-math/src/field.rs; the batched driver of
neo-reductions/src/sumcheck.rs:350-456 folds every oracle of a group with the same challenge)."""
import pytest

Q = 0xFFFFFFFF00000001


def kadd(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def kmul(a, b):
    return ((a[0] * b[0] + 7 * a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def at(vals, x):
    """The polynomial through (j, vals[j]), j = 0..d, evaluated at x in K."""
    d = len(vals) - 1
    acc = (0, 0)
    for j in range(d + 1):
        num, den = (1, 0), 1
        for m in range(d + 1):
            if m != j:
                num = kmul(num, ((x[0] - m) % Q, x[1]))
                den = den * (j - m) % Q
        inv = pow(den, Q - 2, Q)
        acc = kadd(acc, kmul(vals[j], kmul(num, (inv, 0))))
    return acc


def test_interpolation_reproduces_nodes_and_a_cubic():
    vals = [(5, 1), (7, 0), (Q - 2, 3), (11, 4)]
    for j, v in enumerate(vals):
        assert at(vals, (j, 0)) == v
    # p(x) = 3 + 2x + x^3 in F_q (u-part 0): from its values at 0..3, p(10) = 1023
    cubic = [((3 + 2 * x + x ** 3) % Q, 0) for x in range(4)]
    assert at(cubic, (10, 0)) == (1023, 0)


@pytest.mark.parametrize("ell", [5, 6])
def test_chunk_sumcheck_chains(oracle, ell):
    from oracle import pyoracle
    from nightstream_amd.sidecar import chunk_lanes, chunk_rounds
    pyoracle.build()
    lanes = chunk_lanes(11 + ell, ell)
    rounds = chunk_rounds(pyoracle.TsOracle, lanes, ell, seed=99 + ell)
    assert len(rounds) == 30
    with_rounds = 0
    for label, claim, per, chals in rounds:
        assert len(per) == len(chals)
        # the Twist read and write checks' constructors return no claim: their chain starts at
        # the first round's own sum
        running = kadd(per[0][0], per[0][1]) if label.endswith(("/read", "/write")) else claim
        for vals, r in zip(per, chals):
            assert kadd(vals[0], vals[1]) == running, label
            running = at(vals, r)
        with_rounds += bool(per)
    assert with_rounds == 30
