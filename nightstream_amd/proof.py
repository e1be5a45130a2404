"""Canonical proof encoding of a CCS-only ShardProof and its reader.

The reference's ShardProof has no wire format (neo-fold/src/shard_proof_types.rs:166 derives no
serde).  Both the MI355X prover and the CPU restatement emit the same little-endian u64 word
stream, field-for-field in the order of the Rust structs, so proofs compare bit-for-bit:

    tag "NPROPRF1", n_steps, then per StepProof:
      FoldStep.ccs_out: count, MeInstance*
      FoldStep.ccs_proof (PiCcsProof, optimized_engine/mod.rs:57-80):
        sumcheck_rounds: count, (len, K*len)*; sc_initial_sum: K;
        sumcheck_challenges: count, K*; challenges_public: alpha, beta_a, beta_r (count, K*), gamma K;
        sumcheck_final: K; header_digest: 4 words (32 bytes)
      FoldStep.rlc_rhos: count, 54*54 F_q each; rlc_parent: MeInstance; dec_children: count, MeInstance*
      mem (MemSidecarProof, shard_proof_types.rs:120-125):
        cpu_me_claims_val: count, MeInstance*
        shout_addr_pre: claimed_sums (count, K*), active_mask, round_polys (count, rounds*), r_addr (count, K*)
        proofs: count, then per instance a tag (0 Twist, 1 Shout) and
          BatchedAddrProof: claimed_sums (count, K*), round_polys (count, rounds*), r_addr (count, K*);
          Twist only, val_eval: 0 (None) | 1, claimed_inc_sum_lt K, rounds_lt, claimed_inc_sum_total K,
            rounds_total, claimed_prev_inc_sum_total (0 | 1 K), rounds_prev_total (0 | 1 rounds)
        (rounds = count, (len, K*)*); a CCS-only step's mem is 6 zero words
      batched_time: claimed_sums (count, K*), degree_bounds (count, u64*), labels (count, (len, bytes)*),
                    round_polys (count, rounds*)
      val_fold: 0 (None) | 1, rlc_rhos, rlc_parent, dec_children (as in FoldStep)
    ShardProof.output_proof: 0 (None) | 1, output_sc.round_polys (rounds; 4 coefficients each)
    ShardObligations: main (count, MeInstance*), val (count, MeInstance*)
  MeInstance (neo-ccs/src/relations.rs): c (kappa*D words, col-major), m_in, X (D*m_in),
    r (count, K*), y (t, (64, K*64)*), y_scalars (t, K*), fold_digest (4 words),
    c_step_coords.len() (0), u_offset (0), u_len (0).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

D = 54
KAPPA = 16
TAG = 0x314652504F52504E


@dataclass
class MeInstance:
    c: np.ndarray            # kappa*D
    m_in: int
    X: np.ndarray            # D x m_in
    r: list                  # [(c0, c1)]
    y: list                  # t x 64 [(c0, c1)]
    y_scalars: list
    fold_digest: bytes


@dataclass
class TwistAddrVal:
    claimed_sums: list
    round_polys: list        # [claim][round] -> coefficients
    r_addr: list
    val_eval: dict | None    # lt / total / prev_total: (claimed_sum, rounds)


@dataclass
class MemSidecar:
    cpu_me_claims_val: list
    shout_claimed_sums: list
    shout_active_mask: int
    shout_round_polys: list
    shout_r_addr: list
    twist: list              # TwistAddrVal per Twist instance
    n_shout_proofs: int


@dataclass
class StepProof:
    ccs_out: list
    sumcheck_rounds: list
    sc_initial_sum: tuple
    sumcheck_challenges: list
    alpha: list
    beta_a: list
    beta_r: list
    gamma: tuple
    sumcheck_final: tuple
    header_digest: bytes
    rlc_rhos: list
    rlc_parent: MeInstance
    dec_children: list
    batched_claimed_sums: list = field(default_factory=list)
    batched_degree_bounds: list = field(default_factory=list)
    batched_labels: list = field(default_factory=list)
    batched_round_polys: list = field(default_factory=list)
    mem: MemSidecar | None = None
    val_fold: tuple | None = None  # (rhos, parent, children)


@dataclass
class ShardProof:
    steps: list
    final_main: list
    final_val: list = field(default_factory=list)
    output_rounds: list | None = None  # OutputSumcheckProof.round_polys, when output binding is on


class _Reader:
    def __init__(self, data: bytes):
        if len(data) % 8:
            raise ValueError("proof length is not a multiple of 8")
        self.w = np.frombuffer(data, dtype="<u8")
        self.i = 0

    def u(self) -> int:
        v = int(self.w[self.i])
        self.i += 1
        return v

    def k(self) -> tuple:
        return (self.u(), self.u())

    def ks(self) -> list:
        return [self.k() for _ in range(self.u())]

    def words(self, n) -> np.ndarray:
        v = self.w[self.i:self.i + n].copy()
        self.i += n
        return v

    def digest(self) -> bytes:
        return self.words(4).astype("<u8").tobytes()

    def rounds(self) -> list:
        out = []
        for _ in range(self.u()):
            n = self.u()
            out.append([self.k() for _ in range(n)])
        return out

    def addr_proof(self):
        sums = self.ks()
        polys = [self.rounds() for _ in range(self.u())]
        return sums, polys, self.ks()

    def me(self) -> MeInstance:
        c = self.words(KAPPA * D)
        m_in = self.u()
        X = self.words(D * m_in).reshape(D, m_in)
        r = self.ks()
        t = self.u()
        y = []
        for _ in range(t):
            n = self.u()
            y.append([self.k() for _ in range(n)])
        ys = self.ks()
        fd = self.digest()
        if (self.u(), self.u(), self.u()) != (0, 0, 0):
            raise ValueError("unexpected c_step_coords / u_offset / u_len")
        return MeInstance(c, m_in, X, r, y, ys, fd)


def parse(data: bytes) -> ShardProof:
    R = _Reader(data)
    if R.u() != TAG:
        raise ValueError("not a canonical shard proof")
    steps = []
    for _ in range(R.u()):
        ccs_out = [R.me() for _ in range(R.u())]
        rounds = []
        for _ in range(R.u()):
            n = R.u()
            rounds.append([R.k() for _ in range(n)])
        init = R.k()
        chals = R.ks()
        alpha, beta_a, beta_r = R.ks(), R.ks(), R.ks()
        gamma = R.k()
        final = R.k()
        hd = R.digest()
        rhos = [R.words(D * D).reshape(D, D) for _ in range(R.u())]
        parent = R.me()
        children = [R.me() for _ in range(R.u())]
        vclaims = [R.me() for _ in range(R.u())]
        sh_sums = R.ks()
        sh_mask = R.u()
        sh_polys = [R.rounds() for _ in range(R.u())]
        sh_r = R.ks()
        twists, n_shout = [], 0
        for _ in range(R.u()):
            tag = R.u()
            asums, apolys, ar = R.addr_proof()
            if tag == 1:
                n_shout += 1
                continue
            if tag != 0:
                raise ValueError("unknown MemOrLutProof variant")
            ve = None
            if R.u():
                ve = {"lt": (R.k(), R.rounds()), "total": (R.k(), R.rounds())}
                prev_sum = R.k() if R.u() else None
                prev_rounds = R.rounds() if R.u() else None
                if (prev_sum is None) != (prev_rounds is None):
                    raise ValueError("rollover claim without its rounds")
                ve["prev_total"] = None if prev_sum is None else (prev_sum, prev_rounds)
            twists.append(TwistAddrVal(asums, apolys, ar, ve))
        mem = MemSidecar(vclaims, sh_sums, sh_mask, sh_polys, sh_r, twists, n_shout)
        sums = R.ks()
        degs = [R.u() for _ in range(R.u())]
        labels = []
        for _ in range(R.u()):
            n = R.u()
            labels.append(bytes(R.u() for _ in range(n)))
        polys = [R.rounds() for _ in range(R.u())]
        val_fold = None
        if R.u():
            vrhos = [R.words(D * D).reshape(D, D) for _ in range(R.u())]
            vparent = R.me()
            val_fold = (vrhos, vparent, [R.me() for _ in range(R.u())])
        steps.append(StepProof(ccs_out, rounds, init, chals, alpha, beta_a, beta_r, gamma, final, hd, rhos, parent,
                               children, sums, degs, labels, polys, mem, val_fold))
    output_rounds = R.rounds() if R.u() else None
    final_main = [R.me() for _ in range(R.u())]
    final_val = [R.me() for _ in range(R.u())]
    if R.i != len(R.w):
        raise ValueError("trailing words after the proof")
    return ShardProof(steps, final_main, final_val, output_rounds)
