"""Shard workload with Twist/Shout memory over a shared CPU bus (the RISC-V-like trace shard).

A chunk of `c` cycles per fold step.  Per cycle: two register reads (rs1, rs2) and one register
write (rd) over 32 u32 registers; one Shout lookup into a 2^16-entry 8-bit x 8-bit table whose
op (ADD / XOR / AND) is uniform per cycle, with uniform operands, and whose result is the value
written to rd; with probability 0.35 one RAM access to a 2^16-word memory at a Zipf(1.1)
address, 40% of them writes of uniform u32 values.

Memory instances (neo-memory witness types, `witness.rs:23-96`), in bus order:
  Shout 0..2: the ADD, XOR and AND tables (k = 2^16, d = 1, ell = 16), one lookup lane each;
  Twist 0: registers, read port rs1, write port rd (k = 32, ell = 5);
  Twist 1: the same register file mirrored, read port rs2, write port rd (one Twist instance
           has one read and one write port per row, `cpu/bus_layout.rs:11-27`);
  Twist 2: RAM (k = 2^16, ell = 16).
Each step's MemInit is the memory state at the chunk's start (registers: 32 random u32 values
before step 0; RAM: 64 random words before step 0).

The bus layout (`neo-memory/src/cpu/bus_layout.rs:98-191`): Shout columns (addr_bits, has_lookup,
val), then Twist columns (ra_bits, wa_bits, has_read, has_write, wv, rv, inc): 3*18 + 2*15 + 37 =
121 columns, cell (col, j) at z[bus_base + col*c + j], time index t = m_in + j.  The reference's
shared-bus guardrails (`neo-fold/src/memory_sidecar/cpu_bus.rs:157-181, 612-941`) need the CCS to
hold the bus region (m >= 121 c), a canonical padding row (1 - flag) * field = 0 for every gated
bus field of every cycle, and a reference to every bus column except inc outside those rows.  The
CCS is therefore an R1CS with n = m = 128 c (2^19 at c = 2^12):

  z:    [0] = 1 (the public input) | core [1, c) | CPU columns r1, r2, rd, vd, ma, lo (c each) | bus
  rows: [0, c)       the R1CS generator's rows over the core columns (A, B: 3 uniform columns with
                     uniform coefficients; C: one column with the solving coefficient)
        112 per cycle: padding rows  (one - flag) * field = 0
        13 per cycle:  binding rows  r1 = sum 2^b ra_bits(rs1 Twist), r2 = ... (rs2 Twist),
                       rd = sum 2^b wa_bits (both register Twists), vd = wv (both), vd = the
                       lookup result (sum of the three lanes' val), lo = sum 2^b addr_bits (all
                       lanes), exactly one lookup per cycle, has_read + has_write = 2 (both
                       register Twists), ma = sum 2^b (ra_bits + wa_bits) (RAM); and one row
                       s * 1 = s over the read/written values and RAM flags, which puts them in an
                       active constraint as the guardrail asks and constrains nothing more (their
                       values are checked by the Twist read/write arguments)
        the remaining 2c rows are zero.
"""
from __future__ import annotations

import dataclasses

import numpy as np

from .workload import Q, R1CS, rand_balanced_z, rand_field

SHOUT_ELL = 16
REG_ELL = 5
RAM_ELL = 16
N_CPU = 6


@dataclasses.dataclass
class LutInst:
    """LutInstance (neo-memory/src/witness.rs:82-96) with a materialized table, no table_spec."""
    k: int
    d: int
    n_side: int
    steps: int
    ell: int
    table: np.ndarray  # k F_q values


@dataclasses.dataclass
class MemInst:
    """MemInstance (witness.rs:23-36); init = None is MemInit::Zero, else sorted (addr, value)."""
    k: int
    d: int
    n_side: int
    steps: int
    ell: int
    init: list | None


@dataclasses.dataclass
class StepMem:
    luts: list
    mems: list


@dataclasses.dataclass
class BusLayout:
    """build_bus_layout_for_instances (bus_layout.rs:98-191)."""
    m: int
    m_in: int
    chunk: int
    bus_cols: int
    bus_base: int
    shout: list  # (addr_bits range start, ell, has_lookup, val)
    twist: list  # (ra start, wa start, ell, has_read, has_write, wv, rv, inc)

    def cell(self, col: int, j):
        return self.bus_base + col * self.chunk + j


def bus_layout(m: int, m_in: int, chunk: int, shout_ells, twist_ells) -> BusLayout:
    col, shout, twist = 0, [], []
    for e in shout_ells:
        shout.append((col, e, col + e, col + e + 1))
        col += e + 2
    for e in twist_ells:
        twist.append((col, col + e, e, col + 2 * e, col + 2 * e + 1, col + 2 * e + 2, col + 2 * e + 3, col + 2 * e + 4))
        col += 2 * e + 5
    if col * chunk > m:
        raise ValueError(f"bus region {col}*{chunk} exceeds m={m}")
    base = m - col * chunk
    if base < m_in:
        raise ValueError("bus region overlaps the public inputs")
    return BusLayout(m, m_in, chunk, col, base, shout, twist)


def _tables():
    a = np.arange(1 << 16, dtype=np.uint64) & np.uint64(0xFF)
    b = np.arange(1 << 16, dtype=np.uint64) >> np.uint64(8)
    return [a + b, a ^ b, a & b]


def _fq(v) -> np.ndarray:
    """Signed int64 values as F_q elements."""
    v = np.asarray(v, dtype=np.int64)
    return np.where(v >= 0, v.astype(np.uint64), np.uint64(Q) - (-v).astype(np.uint64)).astype(np.uint64)


class _Rows:
    """Triplets (row, col, val) of one matrix, collected in blocks and turned into CSR."""

    def __init__(self):
        self.r, self.c, self.v = [], [], []

    def add(self, rows, cols, vals):
        rows = np.broadcast_to(np.asarray(rows, np.int64), np.shape(cols))
        self.r.append(np.asarray(rows, np.int64).ravel())
        self.c.append(np.asarray(cols, np.int64).ravel())
        self.v.append(np.broadcast_to(np.asarray(vals, np.uint64), np.shape(cols)).ravel())

    def csr(self, n):
        r = np.concatenate(self.r) if self.r else np.zeros(0, np.int64)
        c = np.concatenate(self.c) if self.c else np.zeros(0, np.int64)
        v = np.concatenate(self.v) if self.v else np.zeros(0, np.uint64)
        keep = v != 0
        r, c, v = r[keep], c[keep], v[keep]
        order = np.lexsort((c, r))
        r, c, v = r[order], c[order], v[order]
        if len(r) > 1 and np.any((r[1:] == r[:-1]) & (c[1:] == c[:-1])):
            raise ValueError("duplicate matrix entry")
        ptr = np.zeros(n + 1, np.uint64)
        np.add.at(ptr, r + 1, 1)
        return np.cumsum(ptr).astype(np.uint64), c.astype(np.uint32), v.astype(np.uint64)


def gen_bus_shard(log_chunk: int, nsteps: int, seed: int, with_final: bool = False):
    """The CCS (an R1CS with n = m = 128 * 2^log_chunk), one witness z per step and the step's
    memory instances.  Returns (R1CS, [z], [StepMem], BusLayout), and with with_final also the
    memory after the last step per Twist instance ({mem_idx: u64[k]}, the final_memory_state of
    output binding)."""
    c = 1 << log_chunk
    n = m = 128 * c
    rng = np.random.default_rng(seed)
    lay = bus_layout(m, 1, c, [SHOUT_ELL] * 3, [REG_ELL, REG_ELL, RAM_ELL])
    assert lay.bus_base == 7 * c and lay.bus_cols == 121
    j = np.arange(c, dtype=np.int64)
    one = 0
    r1c, r2c, rdc, vdc, mac, loc = (c * (1 + i) + j for i in range(N_CPU))
    A, B, C = _Rows(), _Rows(), _Rows()
    ONE = np.uint64(1)
    NEG1 = np.uint64(Q - 1)

    # core rows [0, c): the R1CS generator over the core columns (constant across steps)
    zc = rand_balanced_z(rng, c)
    zco = zc.astype(object)
    idx = rng.integers(0, c, size=(2, c, 3))
    for t in range(2):
        while True:
            dup = (idx[t, :, 0] == idx[t, :, 1]) | (idx[t, :, 0] == idx[t, :, 2]) | (idx[t, :, 1] == idx[t, :, 2])
            if not dup.any():
                break
            idx[t][dup] = rng.integers(0, c, size=(int(dup.sum()), 3))
    av = rand_field(rng, 3 * c).reshape(c, 3)
    bv = rand_field(rng, 3 * c).reshape(c, 3)
    A.add(j[:, None], idx[0], av)
    B.add(j[:, None], idx[1], bv)
    az = (av.astype(object) * zco[idx[0]]).sum(axis=1) % Q
    bz = (bv.astype(object) * zco[idx[1]]).sum(axis=1) % Q
    nz_core = np.nonzero(zc[1:] != 0)[0] + 1
    cj = nz_core[rng.integers(0, len(nz_core), size=c)]
    inv = np.array([pow(int(v), Q - 2, Q) for v in zc[cj]], dtype=object)
    C.add(j, cj, np.array((az * bz % Q) * inv % Q, dtype=np.uint64))

    # padding rows: (one - flag) * field = 0, in the guardrail's order (cpu_bus.rs:612-692)
    pairs = []
    for (a0, e, hl, val) in lay.shout:
        pairs.append((hl, val))
        pairs += [(hl, a0 + b) for b in range(e)]
    for (ra0, wa0, e, hr, hw, wv, rv, inc) in lay.twist:
        pairs.append((hr, rv))
        pairs += [(hr, ra0 + b) for b in range(e)]
        pairs.append((hw, wv))
        pairs.append((hw, inc))
        pairs += [(hw, wa0 + b) for b in range(e)]
    assert len(pairs) == 112
    row = c
    for (fl, fd) in pairs:
        rows = row + np.arange(c)
        A.add(rows, np.full(c, one), ONE)
        A.add(rows, lay.cell(fl, j), NEG1)
        B.add(rows, lay.cell(fd, j), ONE)
        row += c

    # binding rows (13 per cycle)
    def lin(cols_vals, rhs_one=0):
        nonlocal row
        rows = row + np.arange(c)
        for col, val in cols_vals:
            A.add(rows, col, val)
        if rhs_one:
            A.add(rows, np.full(c, one), _fq([-rhs_one])[0])
        B.add(rows, np.full(c, one), ONE)
        row += c

    p2 = lambda b: _fq([-(1 << b)])[0]  # noqa: E731
    (s_a, s_x, s_n) = lay.shout
    (ta, tb, tr) = lay.twist
    lin([(r1c, ONE)] + [(lay.cell(ta[0] + b, j), p2(b)) for b in range(REG_ELL)])
    lin([(r2c, ONE)] + [(lay.cell(tb[0] + b, j), p2(b)) for b in range(REG_ELL)])
    lin([(rdc, ONE)] + [(lay.cell(ta[1] + b, j), p2(b)) for b in range(REG_ELL)])
    lin([(rdc, ONE)] + [(lay.cell(tb[1] + b, j), p2(b)) for b in range(REG_ELL)])
    lin([(vdc, ONE), (lay.cell(ta[5], j), NEG1)])
    lin([(vdc, ONE), (lay.cell(tb[5], j), NEG1)])
    lin([(vdc, ONE)] + [(lay.cell(s[3], j), NEG1) for s in lay.shout])
    lin([(loc, ONE)] + [(lay.cell(s[0] + b, j), p2(b)) for s in lay.shout for b in range(SHOUT_ELL)])
    lin([(lay.cell(s[2], j), ONE) for s in lay.shout], rhs_one=1)
    lin([(lay.cell(ta[3], j), ONE), (lay.cell(ta[4], j), ONE)], rhs_one=2)
    lin([(lay.cell(tb[3], j), ONE), (lay.cell(tb[4], j), ONE)], rhs_one=2)
    lin([(mac, ONE)] + [(lay.cell(tr[0] + b, j), p2(b)) for b in range(RAM_ELL)]
        + [(lay.cell(tr[1] + b, j), p2(b)) for b in range(RAM_ELL)])
    rows = row + np.arange(c)
    taut = [lay.cell(ta[6], j), lay.cell(tb[6], j), lay.cell(tr[6], j), lay.cell(tr[5], j), lay.cell(tr[3], j),
            lay.cell(tr[4], j)]
    for col in taut:
        A.add(rows, col, ONE)
        C.add(rows, col, ONE)
    B.add(rows, np.full(c, one), ONE)
    row += c
    assert row == 126 * c

    inst = R1CS(n, m, A.csr(n), B.csr(n), C.csr(n), None)
    tables = _tables()

    # the trace
    regs = rng.integers(0, 1 << 32, size=32).astype(np.int64)
    K_RAM = 1 << RAM_ELL
    ranks = np.arange(1, K_RAM + 1, dtype=np.float64)
    pz = ranks ** -1.1
    pz /= pz.sum()
    zperm = rng.permutation(K_RAM)
    ram = {int(a): int(v) for a, v in zip(rng.choice(K_RAM, 64, replace=False), rng.integers(0, 1 << 32, 64))}
    zs, mems = [], []
    for _ in range(nsteps):
        reg_init = [(a, int(regs[a]) % Q) for a in range(32) if regs[a] != 0]
        ram_init = sorted((a, v % Q) for a, v in ram.items() if v != 0)
        z = np.zeros(m, np.uint64)
        z[0] = 1
        z[1:c] = zc[1:]
        cells = np.zeros((lay.bus_cols, c), np.int64)
        rs1 = rng.integers(0, 32, c)
        rs2 = rng.integers(0, 32, c)
        rd = rng.integers(0, 32, c)
        op = rng.integers(0, 3, c)
        opa = rng.integers(0, 256, c)
        opb = rng.integers(0, 256, c)
        lo = opa | (opb << 8)
        acc = rng.random(c) < 0.35
        isw = acc & (rng.random(c) < 0.4)
        addr = zperm[rng.choice(K_RAM, size=c, p=pz)]
        wvals = rng.integers(0, 1 << 32, c)
        cpu = np.zeros((N_CPU, c), np.int64)
        for t in range(c):
            res = int(tables[op[t]][lo[t]])
            s = lay.shout[op[t]]
            for b in range(SHOUT_ELL):
                cells[s[0] + b, t] = (lo[t] >> b) & 1
            cells[s[2], t] = 1
            cells[s[3], t] = res
            for (tw, rs) in ((ta, rs1[t]), (tb, rs2[t])):
                ra0, wa0, e, hr, hw, wv, rv, inc = tw
                for b in range(e):
                    cells[ra0 + b, t] = (rs >> b) & 1
                    cells[wa0 + b, t] = (rd[t] >> b) & 1
                cells[hr, t] = 1
                cells[hw, t] = 1
                cells[rv, t] = regs[rs]
                cells[wv, t] = res
                cells[inc, t] = res - regs[rd[t]]
            ra0, wa0, e, hr, hw, wv, rv, inc = tr
            ma = 0
            if acc[t]:
                a = int(addr[t])
                ma = a
                old = ram.get(a, 0)
                if isw[t]:
                    for b in range(e):
                        cells[wa0 + b, t] = (a >> b) & 1
                    cells[hw, t] = 1
                    cells[wv, t] = wvals[t]
                    cells[inc, t] = int(wvals[t]) - old
                    ram[a] = int(wvals[t])
                else:
                    for b in range(e):
                        cells[ra0 + b, t] = (a >> b) & 1
                    cells[hr, t] = 1
                    cells[rv, t] = old
            cpu[:, t] = [rs1[t], rs2[t], rd[t], res, ma, lo[t]]
            regs[rd[t]] = res
        z[c:7 * c] = _fq(cpu.reshape(-1))
        z[lay.bus_base:] = _fq(cells.reshape(-1))
        zs.append(z)
        luts = [LutInst(1 << 16, 1, 1 << 16, c, SHOUT_ELL, tb_.astype(np.uint64)) for tb_ in tables]
        mems.append(StepMem(luts, [MemInst(32, 1, 32, c, REG_ELL, reg_init),
                                   MemInst(32, 1, 32, c, REG_ELL, list(reg_init)),
                                   MemInst(K_RAM, 1, K_RAM, c, RAM_ELL, ram_init)]))
    inst.z = zs[0]
    if not with_final:
        return inst, zs, mems, lay
    reg_final = _fq(regs)
    ram_final = np.zeros(K_RAM, np.uint64)
    for a, v in ram.items():
        ram_final[a] = v % Q
    return inst, zs, mems, lay, {0: reg_final, 1: reg_final.copy(), 2: ram_final}


def r1cs_satisfied(inst: R1CS, z: np.ndarray) -> bool:
    """(A z) o (B z) == C z over F_q, row by row (vectorised with Python integers)."""
    zo = z.astype(object)

    def mv(mat):
        ptr, col, val = mat
        prod = val.astype(object) * zo[col.astype(np.int64)]
        out = np.zeros(inst.n, dtype=object)
        rows = np.repeat(np.arange(inst.n), np.diff(ptr.astype(np.int64)))
        np.add.at(out, rows, prod)
        return out % Q

    return bool(np.all((mv(inst.a) * mv(inst.b)) % Q == mv(inst.c)))
