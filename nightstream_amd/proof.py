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
      mem (MemSidecarProof, empty for CCS-only steps): 6 zero words
      batched_time: claimed_sums (count, K*), degree_bounds (count, u64*), labels (count, (len, bytes)*),
                    round_polys (count, (count, (len, K*)*)*)
      val_fold: 0 (None)
    final obligations (main lane): count, MeInstance*
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


@dataclass
class ShardProof:
    steps: list
    final_main: list


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
        if any(R.u() for _ in range(6)):
            raise ValueError("non-empty memory sidecar in a CCS-only proof")
        sums = R.ks()
        degs = [R.u() for _ in range(R.u())]
        labels = []
        for _ in range(R.u()):
            n = R.u()
            labels.append(bytes(R.u() for _ in range(n)))
        polys = []
        for _ in range(R.u()):
            per = []
            for _ in range(R.u()):
                n = R.u()
                per.append([R.k() for _ in range(n)])
            polys.append(per)
        if R.u() != 0:
            raise ValueError("val_fold present in a CCS-only proof")
        steps.append(StepProof(ccs_out, rounds, init, chals, alpha, beta_a, beta_r, gamma, final, hd, rhos, parent,
                               children, sums, degs, labels, polys))
    final_main = [R.me() for _ in range(R.u())]
    if R.i != len(R.w):
        raise ValueError("trailing words after the proof")
    return ShardProof(steps, final_main)
