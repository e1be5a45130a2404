"""Synthetic workloads for the Neo folding hot path.

R1CS generator (the benchmark workloads' description): n constraints x m variables with n == m.
A and B have 3 nonzeros per row at distinct uniform columns with uniform nonzero Goldilocks
coefficients; C has one nonzero per row at a uniform column j != 0 with coefficient
(Az)_i (Bz)_i / z_j, so A z o B z = C z holds.  z[0] = 1 is the single public input and the
other entries are uniform in the balanced range |z| <= 2^54 - 1: with b = 2 and D = 54 the
balanced decomposition (neo-ajtai/src/decomp.rs:118-158) represents exactly that range, so a
z entry uniform in all of F_q would not recompose (neo-ajtai/tests/decomp_representability_limit.rs).

The instance is fed unchanged to the MI355X prover and to the CPU restatement.
"""
from __future__ import annotations

import numpy as np

Q = 0xFFFFFFFF00000001
ZMAX = (1 << 54) - 1


def _rand_u64(rng: np.random.Generator, size: int) -> np.ndarray:
    hi = rng.integers(0, 1 << 32, size=size, dtype=np.uint64)
    lo = rng.integers(0, 1 << 32, size=size, dtype=np.uint64)
    return (hi << np.uint64(32)) | lo


def rand_field(rng: np.random.Generator, size: int, nonzero: bool = True) -> np.ndarray:
    """Uniform F_q elements by rejection (x < q, and x != 0 when nonzero)."""
    out = np.empty(size, dtype=np.uint64)
    filled = 0
    while filled < size:
        x = _rand_u64(rng, size - filled)
        ok = x < np.uint64(Q)
        if nonzero:
            ok &= x != 0
        x = x[ok]
        out[filled:filled + len(x)] = x
        filled += len(x)
    return out


def rand_balanced_z(rng: np.random.Generator, m: int) -> np.ndarray:
    """z[0] = 1, z[1:] uniform in [-(2^54-1), 2^54-1] mapped into F_q."""
    mag = rng.integers(-ZMAX, ZMAX + 1, size=m, dtype=np.int64)
    z = np.where(mag >= 0, mag.astype(np.uint64), (np.uint64(Q) - (-mag).astype(np.uint64)))
    z = z.astype(np.uint64)
    z[0] = 1
    return z


class R1CS:
    """CSR matrices A, B, C (n x m) and a satisfying witness z (z[0] = 1, m_in = 1)."""

    def __init__(self, n, m, a, b, c, z):
        self.n, self.m = n, m
        self.a, self.b, self.c = a, b, c  # (rowptr u64[n+1], cols u32[nnz], vals u64[nnz])
        self.z = z
        self.m_in = 1


def gen_r1cs(n: int, seed: int, nsteps: int = 1) -> tuple[R1CS, list[np.ndarray]]:
    """One R1CS instance and the witnesses of `nsteps` fold steps.

    A shard proves many steps against one CCS structure.  C's coefficients are solved from z,
    so a second independent uniform z would not satisfy the same C; every step of the shard
    therefore folds the same satisfying z.  Proving work is independent of witness values
    (exact field arithmetic, digit planes of the same distribution), and every step still
    folds a different accumulator because each step draws fresh challenges.
    """
    m = n
    rng = np.random.default_rng(seed)
    z = rand_balanced_z(rng, m)
    zi = [int(v) for v in z]

    def rand_rows():
        cols = np.empty((n, 3), dtype=np.uint32)
        for i in range(n):
            cols[i] = rng.choice(m, size=3, replace=False)
        cols.sort(axis=1)
        vals = rand_field(rng, 3 * n).reshape(n, 3)
        return cols, vals

    a_cols, a_vals = rand_rows()
    b_cols, b_vals = rand_rows()
    c_cols = np.empty(n, dtype=np.uint32)
    c_vals = np.empty(n, dtype=np.uint64)
    jj = rng.integers(1, m, size=n)
    for i in range(n):
        az = sum(int(a_vals[i, k]) * zi[int(a_cols[i, k])] for k in range(3)) % Q
        bz = sum(int(b_vals[i, k]) * zi[int(b_cols[i, k])] for k in range(3)) % Q
        j = int(jj[i])
        while zi[j] == 0:
            j = 1 + (j % (m - 1))
        c_cols[i] = j
        c_vals[i] = (az * bz % Q) * pow(zi[j], Q - 2, Q) % Q
    rp3 = np.arange(0, 3 * n + 1, 3, dtype=np.uint64)
    rp1 = np.arange(0, n + 1, dtype=np.uint64)
    a = (rp3, a_cols.reshape(-1).copy(), a_vals.reshape(-1).copy())
    b = (rp3, b_cols.reshape(-1).copy(), b_vals.reshape(-1).copy())
    c = (rp1, c_cols, c_vals)
    inst = R1CS(n, m, a, b, c, z)
    return inst, [z.copy() for _ in range(nsteps)]


def gen_r1cs_steps(n: int, seed: int, nsteps: int, free_frac: float = 0.25) -> tuple[R1CS, list[np.ndarray]]:
    """One R1CS instance satisfied by `nsteps` DISTINCT witnesses (one per fold step).

    The generator's distribution per row (A, B: 3 distinct uniform columns with uniform nonzero
    coefficients; C: one column j != 0 with coefficient (Az)_i (Bz)_i / z_j) is drawn over a
    "bound" set of columns that holds z[0] and a random (1 - free_frac) share of the others.  The
    remaining "free" columns appear in no constraint, so every step's witness keeps the bound
    entries and draws its free entries afresh (uniform in |z| <= 2^54 - 1): the steps' digit
    matrices, MCS commitments and accumulators all differ, while one CCS structure holds for the
    whole shard, as in fold_shard_prove (neo-fold/src/shard.rs:1725-1745).  Vectorised, so 2^16
    rows take about a second.
    """
    m = n
    rng = np.random.default_rng(seed)
    perm = rng.permutation(np.arange(1, m))
    nfree = int((m - 1) * free_frac)
    free = np.sort(perm[:nfree])
    bound = np.concatenate([[0], np.sort(perm[nfree:])]).astype(np.int64)
    z = rand_balanced_z(rng, m)
    zo = z.astype(object)
    nb = len(bound)

    def rows3():
        idx = rng.integers(0, nb, size=(n, 3))
        while True:
            dup = (idx[:, 0] == idx[:, 1]) | (idx[:, 0] == idx[:, 2]) | (idx[:, 1] == idx[:, 2])
            if not dup.any():
                break
            idx[dup] = rng.integers(0, nb, size=(int(dup.sum()), 3))
        cols = np.sort(bound[idx], axis=1)
        vals = rand_field(rng, 3 * n).reshape(n, 3)
        return cols.astype(np.uint32), vals

    a_cols, a_vals = rows3()
    b_cols, b_vals = rows3()
    az = (a_vals.astype(object) * zo[a_cols.astype(np.int64)]).sum(axis=1) % Q
    bz = (b_vals.astype(object) * zo[b_cols.astype(np.int64)]).sum(axis=1) % Q
    nonzero_bound = bound[1:][z[bound[1:]] != 0]
    jj = nonzero_bound[rng.integers(0, len(nonzero_bound), size=n)]
    inv = np.array([pow(int(v), Q - 2, Q) for v in z[jj]], dtype=object)
    c_vals = np.array((az * bz % Q) * inv % Q, dtype=np.uint64)
    rp3 = np.arange(0, 3 * n + 1, 3, dtype=np.uint64)
    rp1 = np.arange(0, n + 1, dtype=np.uint64)
    a = (rp3, a_cols.reshape(-1).copy(), a_vals.reshape(-1).copy())
    b = (rp3, b_cols.reshape(-1).copy(), b_vals.reshape(-1).copy())
    c = (rp1, jj.astype(np.uint32), c_vals)
    inst = R1CS(n, m, a, b, c, z)
    zs = [z.copy()]
    for _ in range(1, nsteps):
        zi = z.copy()
        zi[free] = rand_balanced_z(rng, len(free) + 1)[1:]
        zs.append(zi)
    return inst, zs


def check_r1cs(inst: R1CS, z: np.ndarray) -> bool:
    zi = [int(v) for v in z]

    def mv(mat, i):
        rp, ci, v = mat
        return sum(int(v[k]) * zi[int(ci[k])] for k in range(int(rp[i]), int(rp[i + 1]))) % Q

    return all(mv(inst.a, i) * mv(inst.b, i) % Q == mv(inst.c, i) for i in range(inst.n))


# ---------------- Twist / Shout traces ----------------
def _kconst(v):
    return (int(v) % Q, 0)


def kvals(vs) -> np.ndarray:
    """Base-field values as K = F_q^2 elements: an (n, 2) uint64 array (c0 = v mod q, c1 = 0)."""
    out = np.zeros((len(vs), 2), np.uint64)
    if len(vs):
        out[:, 0] = np.array([int(v) % Q for v in vs], dtype=np.uint64)
    return out


def kones(n: int) -> np.ndarray:
    out = np.zeros((n, 2), np.uint64)
    out[:, 0] = 1
    return out


def sparse_bits(addrs, active, ell_addr):
    """Bit columns of addresses at the active times: column b holds ONE where bit b is set."""
    cols = []
    ts = np.nonzero(active)[0]
    for b in range(ell_addr):
        sel = ts[((addrs[ts] >> b) & 1) == 1]
        cols.append((sel.astype(np.uint64), kones(len(sel))))
    return cols


def gen_memory_trace(ell_n: int, ell_addr: int, p_access: float, write_frac: float, seed: int, zipf_s: float = 0.0):
    """Synthetic Twist memory lane over 2^ell_n cycles: each cycle accesses memory with probability
    p_access (a read or, with probability write_frac, a write), at a uniform address or a Zipf(s)
    address when zipf_s > 0; written values are uniform u32.  Returns the sparse columns of the
    Twist oracles (has_read, rv, ra_bits, has_write, wv, inc, wa_bits) plus the initial memory.
    rv is the value read (memory contents at that time); inc = wv - old value."""
    rng = np.random.default_rng(seed)
    T, A = 1 << ell_n, 1 << ell_addr
    if zipf_s > 0:
        ranks = np.arange(1, A + 1, dtype=np.float64)
        p = ranks ** (-zipf_s)
        p /= p.sum()
        perm = rng.permutation(A)
        addrs = perm[rng.choice(A, size=T, p=p)]
    else:
        addrs = rng.integers(0, A, size=T)
    access = rng.random(T) < p_access
    is_w = access & (rng.random(T) < write_frac)
    is_r = access & ~is_w
    init = {int(a): int(v) for a, v in zip(rng.choice(A, size=min(A, 64), replace=False),
                                              rng.integers(0, 1 << 32, size=min(A, 64)))}
    mem = dict(init)
    rv, wv, inc = [], [], []
    for t in range(T):
        a = int(addrs[t])
        if is_r[t]:
            rv.append(mem.get(a, 0))
        if is_w[t]:
            v = int(rng.integers(0, 1 << 32))
            wv.append(v)
            inc.append((v - mem.get(a, 0)) % Q)
            mem[a] = v
    tr = np.nonzero(is_r)[0].astype(np.uint64)
    tw = np.nonzero(is_w)[0].astype(np.uint64)
    return dict(
        ell_n=ell_n, ell_addr=ell_addr,
        has_read=(tr, kones(len(tr))), rv=(tr[np.array(rv, dtype=np.int64) != 0] if len(tr) else tr,
                                           kvals([v for v in rv if v != 0])),
        ra_bits=sparse_bits(addrs, is_r, ell_addr),
        has_write=(tw, kones(len(tw))), wv=(tw[np.array(wv, dtype=np.int64) != 0] if len(tw) else tw,
                                            kvals([v for v in wv if v != 0])),
        inc=(tw[np.array(inc, dtype=object) != 0] if len(tw) else tw, kvals([v for v in inc if v != 0])),
        wa_bits=sparse_bits(addrs, is_w, ell_addr),
        init=sorted((a, (v % Q, 0)) for a, v in init.items()),
    )


def gen_lookup_trace(ell_n: int, seed: int):
    """Shout lane over 2^ell_n cycles into the 2^16-entry table op(a, b) for 8-bit a, b, with the
    op chosen per lookup among ADD / XOR / AND (the address is a | b << 8; the table is the
    selected op's, so three tables and three lookup lanes).  Returns per op: (addr_bits,
    has_lookup, val, table)."""
    rng = np.random.default_rng(seed)
    T = 1 << ell_n
    a = rng.integers(0, 256, size=T)
    b = rng.integers(0, 256, size=T)
    op = rng.integers(0, 3, size=T)
    grid_a = np.arange(1 << 16) & 0xFF
    grid_b = np.arange(1 << 16) >> 8
    tables = [(grid_a + grid_b), (grid_a ^ grid_b), (grid_a & grid_b)]
    lanes = []
    addr = a | (b << 8)
    for o in range(3):
        act = op == o
        ts = np.nonzero(act)[0].astype(np.uint64)
        vals = tables[o][addr[act]]
        nz = vals != 0
        lanes.append(dict(addr_bits=sparse_bits(addr, act, 16), has_lookup=(ts, kones(len(ts))),
                          val=(ts[nz], kvals(vals[nz])), table=kvals(tables[o])))
    return lanes
