"""Algebraic verifier for canonical CCS-only shard proofs (host side, Python integers).

Restates the checks of fold_shard_verify_impl (neo-fold/src/shard.rs:2482-2994) that do not
depend on re-deriving Fiat-Shamir challenges; challenges are read from the proof:
  * claimed initial sum T from the carried ME inputs (paper_exact.rs:825-944);
  * batched time and Ajtai sum-check chains: degree bound, p(0) + p(1) = running sum,
    running <- p(r) (neo-reductions/src/sumcheck.rs:214-303, 459-559);
  * terminal identity: final running sum = rhs_terminal_identity_paper_exact
    (optimized_engine/common.rs:688-829) evaluated from the Pi_CCS ME outputs;
  * Pi_RLC public recomputation of X, y, y_scalars and the commitment mix
    (neo-reductions/src/api.rs:333-453, session.rs:311-319);
  * Pi_DEC public checks: y, X and c recompose from the children in base b = 2
    (api.rs:458-587), and the shared point r.
Replaying the transcript to re-derive alpha, beta, gamma, the round challenges and the rotation
challenges needs the Poseidon2 constants of the reference, which this checkout does not contain.
"""
from __future__ import annotations

Q = 0xFFFFFFFF00000001
D = 54
DPAD = 64


class VerifyError(AssertionError):
    pass


def kadd(a, b):
    return ((a[0] + b[0]) % Q, (a[1] + b[1]) % Q)


def ksub(a, b):
    return ((a[0] - b[0]) % Q, (a[1] - b[1]) % Q)


def kmul(a, b):
    return ((a[0] * b[0] + 7 * a[1] * b[1]) % Q, (a[0] * b[1] + a[1] * b[0]) % Q)


def kscale(a, s):
    return (a[0] * s % Q, a[1] * s % Q)


K0, K1 = (0, 0), (1, 0)


def poly_eval(coeffs, x):
    r = K0
    for c in reversed(coeffs):
        r = kadd(kmul(r, x), c)
    return r


def eq_lin(a, b):
    return kadd(kmul(ksub(K1, a), ksub(K1, b)), kmul(a, b))


def eq_points(p, q):
    acc = K1
    for a, b in zip(p, q):
        acc = kmul(acc, eq_lin(a, b))
    return acc


def chi(bits):
    w = [K1]
    for b in bits:
        nb = ksub(K1, b)
        w = [kmul(v, nb) for v in w] + [kmul(v, b) for v in w]
    return w


def _ring_mul(a, b):
    """a * b mod X^54 + X^27 + 1 over F_q (Python integers)."""
    tmp = [0] * (2 * D - 1)
    for i, ai in enumerate(a):
        if ai:
            for j, bj in enumerate(b):
                tmp[i + j] += ai * bj
    for i in range(2 * D - 2, D - 1, -1):
        t = tmp[i]
        tmp[i] = 0
        tmp[i - D] -= t
        tmp[i - 27] -= t
    return [v % Q for v in tmp[:D]]


def _check(cond, msg):
    if not cond:
        raise VerifyError(msg)


def _initial_sum(acc, alpha, gamma, t):
    k = 1 + len(acc)
    if k < 2:
        return K0
    ca = chi(alpha)
    gk = K1
    for _ in range(k):
        gk = kmul(gk, gamma)
    inner = K0
    for j in range(t):
        for idx, me in enumerate(acc):
            ye = K0
            for rho in range(min(len(ca), len(me.y[j]))):
                ye = kadd(ye, kmul(me.y[j][rho], ca[rho]))
            w = K1
            for _ in range(idx + 1):
                w = kmul(w, gamma)
            for _ in range(j):
                w = kmul(w, gk)
            inner = kadd(inner, kmul(w, ye))
    return kmul(gk, inner)


def _terminal_rhs(f_terms, ccs_out, r_prime, alpha_prime, alpha, beta_a, beta_r, gamma, r_inputs):
    t = len(ccs_out[0].y)
    k = len(ccs_out)
    eq_beta = kmul(eq_points(alpha_prime, beta_a), eq_points(r_prime, beta_r))
    eq_ar = kmul(eq_points(alpha_prime, alpha), eq_points(r_prime, r_inputs)) if r_inputs is not None else K0
    mv = []
    for j in range(t):
        acc, pw = K0, 1
        for rho in range(D):
            acc = kadd(acc, kscale(ccs_out[0].y[j][rho], pw))
            pw = pw * 2 % Q
        mv.append(acc)
    Fp = K0
    for coeff, exps in f_terms:
        term = (coeff % Q, 0)
        for j, e in enumerate(exps):
            for _ in range(e):
                term = kmul(term, mv[j])
        Fp = kadd(Fp, term)
    ca = chi(alpha_prime)
    nc = K0
    g = gamma
    for me in ccs_out:
        ye = K0
        for rho in range(min(len(ca), len(me.y[0]))):
            ye = kadd(ye, kmul(me.y[0][rho], ca[rho]))
        nc = kadd(nc, kmul(g, kmul(ye, ksub(kmul(ye, ye), K1))))
        g = kmul(g, gamma)
    ev = K0
    if r_inputs is not None and k >= 2:
        gk = K1
        for _ in range(k):
            gk = kmul(gk, gamma)
        for j in range(t):
            for i_abs in range(1, k):
                me = ccs_out[i_abs]
                ye = K0
                for rho in range(min(len(ca), len(me.y[j]))):
                    ye = kadd(ye, kmul(me.y[j][rho], ca[rho]))
                w = K1
                for _ in range(i_abs):
                    w = kmul(w, gamma)
                for _ in range(j):
                    w = kmul(w, gk)
                ev = kadd(ev, kmul(w, ye))
    gk = K1
    for _ in range(k):
        gk = kmul(gk, gamma)
    return kadd(kmul(eq_beta, kadd(Fp, nc)), kmul(eq_ar, kmul(gk, ev)))


def verify_shard(proof, f_terms, ell_n: int, d_sc: int = 6, k_dec: int = 12, acc_init=()):
    """Raise VerifyError unless every step's public algebra holds.  f_terms: [(coeff, exps)]."""
    acc = list(acc_init)
    for s, st in enumerate(proof.steps):
        t = len(st.ccs_out[0].y)
        k = len(st.ccs_out)
        _check(k == 1 + len(acc), f"step {s}: |ccs_out| = {k}, expected {1 + len(acc)}")
        T = _initial_sum(acc, st.alpha, st.gamma, t)
        _check(T == st.sc_initial_sum, f"step {s}: claimed initial sum mismatch")
        _check(st.batched_claimed_sums == [T], f"step {s}: batched claimed sum mismatch")
        _check(st.batched_degree_bounds == [d_sc] and st.batched_labels == [b"ccs/time"], f"step {s}: claim metadata")
        rounds, chals = st.sumcheck_rounds, st.sumcheck_challenges
        _check(len(rounds) == len(chals) == ell_n + 6, f"step {s}: round count")
        _check(st.batched_round_polys == [rounds[:ell_n]], f"step {s}: batched round polys differ from Pi_CCS rounds")
        running = T
        for rd, (poly, r) in enumerate(zip(rounds, chals)):
            _check(len(poly) <= d_sc + 1, f"step {s} round {rd}: degree bound")
            _check(kadd(poly_eval(poly, K0), poly_eval(poly, K1)) == running, f"step {s} round {rd}: p(0)+p(1)")
            running = poly_eval(poly, r)
        _check(running == st.sumcheck_final, f"step {s}: sumcheck_final")
        r_prime, alpha_prime = chals[:ell_n], chals[ell_n:]
        r_in = acc[0].r if acc else None
        rhs = _terminal_rhs(f_terms, st.ccs_out, r_prime, alpha_prime, st.alpha, st.beta_a, st.beta_r, st.gamma, r_in)
        _check(rhs == running, f"step {s}: terminal identity failed")
        for i, me in enumerate(st.ccs_out):
            _check(me.r == r_prime, f"step {s}: ccs_out[{i}].r != r'")
            _check(me.fold_digest == st.header_digest, f"step {s}: fold digest")
            for j in range(t):
                sc, pw = K0, 1
                for rho in range(D):
                    sc = kadd(sc, kscale(me.y[j][rho], pw))
                    pw = pw * 2 % Q
                _check(sc == me.y_scalars[j], f"step {s}: ccs_out[{i}].y_scalars[{j}]")
            if i >= 1:
                _check((me.c == acc[i - 1].c).all() and (me.X == acc[i - 1].X).all(), f"step {s}: carried c/X")
        # ---- Pi_RLC ----
        par = st.rlc_parent
        _check(len(st.rlc_rhos) == k, f"step {s}: |rhos|")
        if k == 1:
            _check((st.rlc_rhos[0] == [[int(a == b) for b in range(D)] for a in range(D)]).all(), "k=1 rho must be I")
            exp_y, exp_X, exp_c = st.ccs_out[0].y, st.ccs_out[0].X, st.ccs_out[0].c
        else:
            exp_y = [[K0] * DPAD for _ in range(t)]
            exp_X = [[0] * par.m_in for _ in range(D)]
            exp_c = [0] * (16 * D)
            for i, (rho, me) in enumerate(zip(st.rlc_rhos, st.ccs_out)):
                R = [[int(v) for v in row] for row in rho]
                a = [R[r][0] for r in range(D)]
                bal = [v if v <= Q // 2 else v - Q for v in a]
                _check(all(-2 <= v <= 2 for v in bal), f"step {s}: rho_{i} coefficients outside [-2, 2]")
                # rho = rot(a): column j = X^j * a mod Phi_81
                col = a[:]
                for j in range(D):
                    _check([R[r][j] for r in range(D)] == col, f"step {s}: rho_{i} is not rot(a)")
                    last = col[D - 1]
                    col = [(-last) % Q] + col[:D - 1]
                    col[27] = (col[27] - last) % Q
                for j in range(t):
                    for rr in range(D):
                        acc_k = K0
                        for kk in range(D):
                            if R[rr][kk]:
                                acc_k = kadd(acc_k, kscale(me.y[j][kk], R[rr][kk]))
                        exp_y[j][rr] = kadd(exp_y[j][rr], acc_k)
                for rr in range(D):
                    for c in range(par.m_in):
                        exp_X[rr][c] = (exp_X[rr][c] + sum(R[rr][kk] * int(me.X[kk][c]) for kk in range(D))) % Q
                for colc in range(16):
                    prod = _ring_mul(a, [int(v) for v in me.c[colc * D:(colc + 1) * D]])
                    for q in range(D):
                        exp_c[colc * D + q] = (exp_c[colc * D + q] + prod[q]) % Q
        _check([list(r) for r in par.y] == [list(r) for r in exp_y], f"step {s}: RLC y mismatch")
        _check([[int(v) for v in row] for row in par.X] == [[int(v) for v in row] for row in exp_X],
               f"step {s}: RLC X mismatch")
        _check([int(v) for v in par.c] == [int(v) for v in exp_c], f"step {s}: RLC commitment mismatch")
        _check(par.r == r_prime, f"step {s}: RLC r")
        # ---- Pi_DEC ----
        ch = st.dec_children
        _check(len(ch) == k_dec, f"step {s}: |children|")
        for j in range(t):
            for q in range(DPAD):
                lhs, pw = K0, 1
                for c in ch:
                    lhs = kadd(lhs, kscale(c.y[j][q], pw))
                    pw = pw * 2 % Q
                _check(lhs == par.y[j][q], f"step {s}: DEC y recomposition")
        for rr in range(D):
            for c in range(par.m_in):
                lhs = sum(int(ch[i].X[rr][c]) << i for i in range(k_dec)) % Q
                _check(lhs == int(par.X[rr][c]), f"step {s}: DEC X recomposition")
        for q in range(16 * D):
            lhs = sum(int(ch[i].c[q]) << i for i in range(k_dec)) % Q
            _check(lhs == int(par.c[q]), f"step {s}: DEC commitment recomposition")
        for c in ch:
            _check(c.r == par.r and c.fold_digest == par.fold_digest, f"step {s}: child r / digest")
        for i, me in enumerate([par] + ch):
            for j in range(t):
                sc, pw = K0, 1
                for rho in range(D):
                    sc = kadd(sc, kscale(me.y[j][rho], pw))
                    pw = pw * 2 % Q
                _check(sc == me.y_scalars[j], f"step {s}: RLC/DEC output {i} y_scalars[{j}]")
        acc = ch
    _check(len(proof.final_main) == len(acc), "final obligations count")
    for a, b in zip(proof.final_main, acc):
        _check((a.c == b.c).all() and (a.X == b.X).all() and a.y == b.y and a.r == b.r
               and a.y_scalars == b.y_scalars and a.fold_digest == b.fold_digest,
               "final obligations differ from last children")
    return True


R1CS_F = [(1, [0, 1, 1, 0]), (Q - 1, [0, 0, 0, 1])]
