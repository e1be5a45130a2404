"""Pi_CCS row-phase oracle of the CPU restatement against a brute-force sum of the paper's Q.

Q is evaluated literally as in neo-reductions/src/engines/paper_exact_engine/paper_exact.rs:136-335
(q_at_point_paper_exact): for X = (X_a, X_r),
    Q = eq(X, beta) * (F(X_r) + sum_i gamma^i NC_i(X))
        + gamma^k * sum_{j, i >= 2} gamma^{i-1} (gamma^k)^{j-1} eq(X, (alpha, r)) Ztilde_i M_j^T chi_{X_r}(X_a),
with NC_i = N(Ztilde_i M_1^T chi_{X_r}(X_a)), N(y) = (y + 1) y (y - 1), and every matrix/vector
term taken as its multilinear extension in X_r.  The round polynomial of row round j is
sum over the Boolean rest of Q(r_0..r_{j-1}, X, rest, X_a); the test compares it with the optimized
streaming oracle (optimized_engine/oracle.rs:82-503, 1047-1178) at 7 nodes and a random point in K,
folds with a random challenge and repeats for every row round.  It also checks the claimed initial
sum against the ME inputs' y (paper_exact.rs:825-944) and precompute_for_r against a dense product.
"""
import numpy as np
import pytest

from kmath import (D, K0, K1, Q, chi, eq_points, from_arr, kadd, kmul, kpow, kscale, ksub, mle, poly_eval,
                   rand_k, to_arr)

ELL_D = 6


def _dense(csr, n, m):
    rp, ci, v = csr
    M = [[0] * m for _ in range(n)]
    for r in range(n):
        for k in range(int(rp[r]), int(rp[r + 1])):
            M[r][int(ci[k])] = (M[r][int(ci[k])] + int(v[k])) % Q
    return M


class Brute:
    def __init__(self, inst, Zs, alpha, beta_a, beta_r, gamma, r_in):
        n, m = inst.n, inst.m
        self.n, self.m = n, m
        self.mats = [[[int(r == c) for c in range(m)] for r in range(n)],
                     _dense(inst.a, n, m), _dense(inst.b, n, m), _dense(inst.c, n, m)]
        self.Zs = [np.asarray(Z, np.int64).reshape(D, m) for Z in Zs]
        self.k = len(Zs)
        self.alpha, self.beta_a, self.beta_r, self.gamma, self.r_in = alpha, beta_a, beta_r, gamma, r_in
        z1 = [sum(int(self.Zs[0][rho, c]) << rho for rho in range(D)) % Q for c in range(m)]
        self.Mz = [[sum(M[r][c] * z1[c] for c in range(m)) % Q for r in range(n)] for M in self.mats]
        # (Z_i M_j^T)[rho][row] = sum_c Z_i[rho, c] M_j[row, c]
        self.ZM = [[[[sum(int(Z[rho, c]) * M[r][c] for c in range(m)) % Q for r in range(n)]
                     for rho in range(D)] for M in self.mats] for Z in self.Zs]
        self.chi_ba = chi(beta_a)
        self.chi_al = chi(alpha)
        self.gk = kpow(gamma, self.k)

    def q_sum_over_rho(self, xr):
        cr = chi(xr)

        def ev(table):
            acc = K0
            for v, c in zip(table, cr):
                if v:
                    acc = kadd(acc, kscale(c, v % Q))
            return acc

        fv = [ev(t) for t in self.Mz]
        F = ksub(kmul(fv[1], fv[2]), fv[3])
        eqr = eq_points(xr, self.beta_r)
        eq_in = eq_points(xr, self.r_in) if self.r_in is not None else K0
        total = K0
        for rho in range(1 << ELL_D):
            nc = K0
            g = self.gamma
            for i in range(self.k):
                if rho < D:
                    y = ev([int(v) for v in self.Zs[i][rho]])
                    ny = kmul(kmul(kadd(y, K1), y), ksub(y, K1))
                    nc = kadd(nc, kmul(g, ny))
                g = kmul(g, self.gamma)
            acc = kmul(kmul(self.chi_ba[rho], eqr), kadd(F, nc))
            if self.r_in is not None and self.k >= 2 and rho < D:
                inner = K0
                for j in range(4):
                    for i in range(1, self.k):
                        w = kmul(kpow(self.gamma, i), kpow(self.gk, j))
                        inner = kadd(inner, kmul(w, ev(self.ZM[i][j][rho])))
                acc = kadd(acc, kmul(self.gk, kmul(kmul(self.chi_al[rho], eq_in), inner)))
            total = kadd(total, acc)
        return total

    def round_poly_at(self, prefix, x, ell_n):
        rest = ell_n - len(prefix) - 1
        acc = K0
        for mask in range(1 << rest):
            bits = [((mask >> b) & 1, 0) for b in range(rest)]
            acc = kadd(acc, self.q_sum_over_rho(list(prefix) + [x] + bits))
        return acc


def _setup(oracle, k_total, seed, n=8):
    from nightstream_amd.workload import gen_r1cs
    rng = np.random.default_rng(seed)
    inst, zs = gen_r1cs(n, seed, 1)
    ccs = oracle.Ccs(inst)
    Zs = [oracle.decomp(zs[0])] + [rng.integers(-1, 2, size=(D, n)).astype(np.int8) for _ in range(k_total - 1)]
    ell_n = n.bit_length() - 1
    alpha = [rand_k(rng) for _ in range(ELL_D)]
    beta_a = [rand_k(rng) for _ in range(ELL_D)]
    beta_r = [rand_k(rng) for _ in range(ell_n)]
    gamma = rand_k(rng)
    r_in = [rand_k(rng) for _ in range(ell_n)] if k_total > 1 else None
    return inst, ccs, Zs, ell_n, alpha, beta_a, beta_r, gamma, r_in, rng


@pytest.mark.parametrize("k_total,seed", [(1, 31), (3, 32)])
def test_row_rounds_match_paper_exact_q(oracle, k_total, seed):
    inst, ccs, Zs, ell_n, alpha, beta_a, beta_r, gamma, r_in, rng = _setup(oracle, k_total, seed)
    st = oracle.RowState(ccs, Zs, to_arr(alpha), to_arr(beta_a), to_arr(beta_r), gamma, ELL_D, ell_n,
                         None if r_in is None else to_arr(r_in))
    bf = Brute(inst, Zs, alpha, beta_a, beta_r, gamma, r_in)
    prefix = []
    running = None
    for rd in range(ell_n):
        coeffs = from_arr(st.coeffs())
        assert len(coeffs) <= 7
        xs = [(v, 0) for v in range(7)] + [rand_k(rng)]
        got = from_arr(st.evals(to_arr(xs)))
        for x, g in zip(xs, got):
            assert g == poly_eval(coeffs, x)
            assert g == bf.round_poly_at(prefix, x, ell_n), f"round {rd} at {x}"
        s01 = kadd(got[0], got[1])
        if running is not None:
            assert s01 == running
        r = rand_k(rng)
        running = poly_eval(coeffs, r)
        st.fold(r)
        prefix.append(r)
    # fully folded tables are the multilinear evaluations at r'
    assert from_arr(st.export(0)) == [eq_points(prefix, beta_r)]
    for i in range(k_total):
        nc = from_arr(st.export(4, i))
        assert nc == [mle([int(v) for v in Zs[i][rho]], prefix) for rho in range(D)]


def test_initial_sum_matches_me_inputs(oracle):
    inst, ccs, Zs, ell_n, alpha, beta_a, beta_r, gamma, r_in, rng = _setup(oracle, 3, 33)
    st = oracle.RowState(ccs, Zs, to_arr(alpha), to_arr(beta_a), to_arr(beta_r), gamma, ELL_D, ell_n, to_arr(r_in))
    g = from_arr(st.evals(to_arr([K0, K1])))
    # T = gamma^k * sum_{j, i >= 2} gamma^{i-1} (gamma^k)^{j-1} ytilde_ij(alpha), y_ij = Z_i M_j^T chi_{r_in}
    y = oracle.precompute_y(ccs, Zs, to_arr(r_in), ell_n)
    k = len(Zs)
    gk = kpow(gamma, k)
    ca = chi(alpha)
    T = K0
    for j in range(4):
        for i in range(1, k):
            yt = K0
            for rho in range(D):
                yt = kadd(yt, kmul((int(y[i, j, rho, 0]), int(y[i, j, rho, 1])), ca[rho]))
            T = kadd(T, kmul(kmul(kpow(gamma, i), kpow(gk, j)), yt))
    T = kmul(gk, T)
    assert kadd(g[0], g[1]) == T


def test_precompute_y_matches_dense_product(oracle):
    inst, ccs, Zs, ell_n, alpha, beta_a, beta_r, gamma, r_in, rng = _setup(oracle, 3, 34, n=16)
    r = [rand_k(rng) for _ in range(ell_n)]
    y = oracle.precompute_y(ccs, Zs, to_arr(r), ell_n)
    cr = chi(r)
    mats = [None, _dense(inst.a, inst.n, inst.m), _dense(inst.b, inst.n, inst.m), _dense(inst.c, inst.n, inst.m)]
    for i, Z in enumerate(Zs):
        for j in range(4):
            if j == 0:
                v = cr[:inst.m]
            else:
                v = [K0] * inst.m
                for row in range(inst.n):
                    for c in range(inst.m):
                        if mats[j][row][c]:
                            v[c] = kadd(v[c], kscale(cr[row], mats[j][row][c]))
            for rho in range(D):
                acc = K0
                for c in range(inst.m):
                    d = int(Z[rho, c])
                    if d:
                        acc = kadd(acc, v[c] if d > 0 else ksub(K0, v[c]))
                assert (int(y[i, j, rho, 0]), int(y[i, j, rho, 1])) == acc


def test_interpolate_recovers_polynomial(oracle):
    rng = np.random.default_rng(35)
    coeffs = [rand_k(rng) for _ in range(7)]
    xs = [(v, 0) for v in range(7)]
    ys = [poly_eval(coeffs, x) for x in xs]
    out = np.zeros(14, np.uint64)
    oracle.lib().or_interpolate(to_arr(xs), to_arr(ys), 7, out)
    assert from_arr(out) == coeffs
