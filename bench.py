#!/usr/bin/env python3
"""Fold-steps/s of the MI355X Neo prover on the 2^16-row R1CS shard.

A "step" is one fold step (neo-fold/src/shard.rs:1803-2273: Pi_CCS sum-check, Ajtai rounds, ME
outputs, Pi_RLC, Pi_DEC with 12 child commitments) including that step's MCS commitment.  The
timed region proves one shard of --steps fold steps (default 16, n = m = 2^16: the
"plain method at scale on one GPU" workload) with the step witnesses already resident in HBM.

Printed JSON line: metric/value/unit plus
  roofline     -- the kernel class with the largest summed duration, timed live by start/stop events
                  that hipExtLaunchKernelGGL records on the prover's stream over the timed region,
                  against MI355X peaks; traffic from the committed FETCH_SIZE/WRITE_SIZE summary;
  cpu_baseline -- the CPU restatement (oracle/, "port") timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PP_SEED = bytes([0x4E, 0x53] + [0x07] * 30)
HBM_PEAK_GBS = 8000.0
INT8_PEAK_TOPS = 5000.0
KT = ["round0_eval", "round_eval", "table_fold", "ajtai_gemm", "ydigits_gemm", "rlc_zmix", "dec_split", "round1_eval",
      "round_fused"]
KERNEL_NAMES = {
    # rounds 0-2: the digit-monomial MFMA kernel, then the E-sum / f / Eval launch (shared by the
    # three rounds); below the MFMA path's sizes, k_round_eval<0>, <1> and <2>
    "round0_eval": ("k_digit_monomials<0>", "k_round_eval<6>"),
    "round1_eval": ("k_digit_monomials<1>", "k_round_eval<6>"),
    "round_eval": ("k_digit_monomials8", "k_round_eval<6>", "k_round_eval<2>"),
    "table_fold": ("k_fold_k", "k_fold_digits3"),
    "ajtai_gemm": ("k_ajtai_gemm32p", "k_ajtai_gemm32", "k_ajtai_gemm"),
    "ydigits_gemm": ("k_ydigits_gemm_lds<4>", "k_ydigits_gemm<4>"),
    "rlc_zmix": "k_rowcontract<4>",
    "dec_split": "k_dec_split",
    "round_fused": "k_round_fused",
}
D = 54


def algorithmic_work(n: int, steps: int, nw_mcs: int):
    """Algorithmic bytes (or int8 ops) per kernel class summed over one shard of `steps`.

    Pi_CCS tables per row at round r (length L = n >> r): k*54 NC entries + nsc scalar tables
    (eq_beta, 3 f-var tables, plus eq_rin and Eval once k > 1), 16 B per K entry.  Rounds 0-2
    read the int8 digit matrices (1 B per digit, all n digits of a row each) instead of NC K
    entries; from n = 512 the fold after round 2 writes the NC tables of length n/8 from the
    digits (8 digits per entry) and round 3 reads them unfused; below, the fold after round 1
    writes them at length n/4 and round 2 reads them (K tables).  Every later round is one fused
    launch that reads the previous round's NC, f-var and Eval tables (2L entries) and writes the
    folded ones (L entries), and reads the chi tables of the later variables (L/2 entries, one for
    eq_beta and one for eq_rin) in place of the eq tables.
    """
    ell_n = n.bit_length() - 1
    out = {c: 0.0 for c in KT}
    for s in range(steps):
        k = 1 if s == 0 else 13
        nsc = 4 if s == 0 else 6
        out["round0_eval"] += k * D * n + nsc * 16 * n
        out["round1_eval"] += k * D * n + nsc * 16 * (n // 2)
        # round 2 from the digits (int8 MFMA, 120 monomials of 8 digits) once n / 8 fills 64-pair
        # blocks; then round 3 unfused on the tables folded from the digits after round 2
        digit2 = n >= 512
        if digit2:
            out["round_eval"] += k * D * n + nsc * 16 * (n >> 2)
            out["round_eval"] += (k * D + nsc) * 16 * (n >> 3)
        else:
            out["round_eval"] += (k * D + nsc) * 16 * (n >> 2)
        nfold, ngt = (3, 1) if s == 0 else (4, 2)
        for r in range(4 if digit2 else 3, ell_n):
            L = n >> r
            out["round_fused"] += (k * D + nfold) * 16 * (2 * L + L) + ngt * 16 * (L // 2)
        # separate folds: scalar tables after rounds 0, 1 (and 2), the NC tables from the digits,
        # and the last round's fold (flushed before y)
        for r in (0, 1, 2) if digit2 else (0, 1):
            L = n >> r
            out["table_fold"] += nsc * 16 * (L + L // 2)
        if digit2:
            out["table_fold"] += k * D * n + k * D * 16 * (n // 8)
        else:
            out["table_fold"] += k * D * n + k * D * 16 * (n // 4)
        out["table_fold"] += (k * D + nsc) * 16 * (2 + 1)
        # y products: k witnesses at the row point + 12 DEC children, t = 4 matrices
        for nw in (k, 12):
            # digits x 64 byte-limb rows of v: 2 (nw 54) 64 n int8 ops on (nw 54 + 64) n bytes, 117
            # ops per byte, under the MI355X's int8 ridge (~625): priced against HBM, by bytes
            out["ydigits_gemm"] += (nw * D + 64) * n
        out["rlc_zmix"] += (k * D * n + D * n * 4) if k > 1 else D * n * 5
        out["dec_split"] += D * n * 4 + 12 * D * n
        out["ajtai_gemm"] += 2.0 * 6912 * D * 12 * n  # 12 DEC children per step
    out["ajtai_gemm"] += 2.0 * 6912 * D * nw_mcs * n  # MCS commitments of all steps (one launch)
    return out


def kernel_bound(cls: str):
    if cls == "ajtai_gemm":
        return "mfma", "TFLOP/s", INT8_PEAK_TOPS
    return "hbm", "GB/s", HBM_PEAK_GBS


def cpu_baseline(inst, zs, cpu_steps: int):
    """The CPU restatement on the first `cpu_steps` witnesses of the timed shard, twice:
    all-cores (OpenMP wherever a loop parallelises) and serial-where-serial (the Pi_RLC Z_mix and
    precompute_for_r loops on one thread, as the reference runs them; Ajtai columns and DEC
    children keep every thread, as the reference's rayon loops do).  Returns the reported baseline
    (serial-where-serial), the all-cores figure and the oracle's proof bytes of those steps."""
    from oracle import pyoracle

    pyoracle.build()
    runs = {}
    proof = None
    for mode in ("serial_where_serial", "all_cores"):
        pyoracle.set_serial_where_serial(mode == "serial_where_serial")
        prover = pyoracle.Prover(inst, PP_SEED)
        t0 = time.perf_counter()
        pr, _ = prover.prove(zs[:cpu_steps])
        wall = time.perf_counter() - t0
        if proof is None:
            proof = pr
        elif pr != proof:
            raise SystemExit("CPU restatement: the two threading modes gave different proofs")
        runs[mode] = (prover.step_seconds, wall)
    pyoracle.set_serial_where_serial(False)
    # steady state: step 1 folds k = 13 claims like every later step; step 0 folds k = 1
    idx = 1 if cpu_steps >= 2 else 0
    secs, wall = runs["serial_where_serial"]
    asecs, awall = runs["all_cores"]
    base = {
        "value": 1.0 / secs[idx],
        "unit": "fold-steps/s",
        "cores": pyoracle.num_threads(),
        "kind": "port",
        "threading": "serial where the reference is serial (Z_mix, precompute_for_r, the row rounds); "
                     "Ajtai columns and DEC children on all cores",
        "sample": (f"CPU restatement (oracle/) of the timed 2^{inst.n.bit_length() - 1}-row shard's first "
                   f"{cpu_steps} steps: fold step {idx} (k = {13 if idx else 1} claims) took {secs[idx]:.2f} s; "
                   f"{cpu_steps} steps + MCS commits {wall:.1f} s wall"),
        "all_cores": {"value": 1.0 / asecs[idx], "cores": pyoracle.num_threads(),
                      "sample": f"same steps with every loop on all cores: step {idx} {asecs[idx]:.2f} s"},
    }
    return base, proof


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=16)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 16)
    ap.add_argument("--cpu-steps", type=int, default=2, help="fold steps of the CPU baseline sample (0 = skip)")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--workload", choices=["shard", "batch", "serve", "twist", "ajtai", "route_a"], default="shard",
                    help="shard: one 2^16-row shard (default); batch: 128 shards of 2^10..2^16 rows, plain path vs "
                         "a pool of concurrent prover contexts; serve: Poisson arrivals with continuous admission")
    ap.add_argument("--workers", type=int, default=16, help="concurrent prover contexts (batch / serve)")
    ap.add_argument("--shared-rounds", type=int, default=0, choices=[0, 1],
                    help="batch / serve: pooled contexts share fused-round launches (1) or not (0)")
    ap.add_argument("--procs", type=int, default=8,
                    help="batch: prover processes sharing the GPU, --workers contexts split evenly (1: one process)")
    ap.add_argument("--proc-hw-queues", type=int, default=1,
                    help="batch: HIP hardware queues per prover process (0: the runtime's default of 4)")
    ap.add_argument("--build-threads", type=int, default=8, help="twist: host threads building a chunk's oracles")
    ap.add_argument("--log-chunk", type=int, default=12, help="route_a: cycles per fold step = 2^log_chunk")
    ap.add_argument("--ajtai-max-log", type=int, default=20, help="ajtai: largest m = 2^k of the sweep")
    ap.add_argument("--shards", type=int, default=128, help="batch: number of shards")
    ap.add_argument("--batch-memory", type=int, default=1, choices=[0, 1],
                    help="batch: every other shard is a Twist/Shout shard (1) or none is (0)")
    ap.add_argument("--serve-memory-frac", type=float, default=0.3,
                    help="serve: fraction of jobs that are Twist/Shout shards")
    ap.add_argument("--mix", type=int, default=1, choices=[0, 1],
                    help="shard: also run a short in-process batch (16 shards, plain vs pooled contexts) and a "
                         "short Poisson serve over the same shards, reported under 'mix'")
    ap.add_argument("--rate-duration", type=float, default=20.0, help="serve: seconds of arrivals per rate")
    ap.add_argument("--rates", type=str, default="", help="serve: comma-separated arrival rates (jobs/s)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="only run 3 plain element-wise copies of 512 MiB (int64 negate: 512 MiB read + 512 MiB "
                         "written each) as the reference kernel for scaling FETCH_SIZE / WRITE_SIZE counters")
    args = ap.parse_args()
    if args.calibrate_copy:
        import torch
        src = torch.arange(1 << 26, dtype=torch.int64, device="cuda:0")
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.neg(src, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_read": src.numel() * 8, "copy_bytes_written": dst.numel() * 8, "launches": 3}))
        return
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 (independent shards "
                                   "shard across GPUs as separate jobs, one process per GPU)"}))
        sys.exit(2)
    if args.steps < 1:
        sys.exit("--steps must be >= 1")
    from nightstream_amd import batch as _batch
    _batch.SHARE_ROUNDS = bool(args.shared_rounds)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "twist":
        return bench_twist(args)
    if args.workload == "serve":
        return bench_serve(args)
    if args.workload == "ajtai":
        return bench_ajtai(args)
    if args.workload == "route_a":
        return bench_route_a(args)

    import nightstream_amd as ns
    from nightstream_amd.workload import gen_r1cs_steps

    if ns.device_count() < 1:
        sys.exit("no GPU visible")
    t_setup = time.perf_counter()
    # one CCS, a distinct witness per fold step (free columns resampled per step)
    nz = max(args.steps, args.warmup, args.cpu_steps, 1)
    inst, zs = gen_r1cs_steps(args.n, args.seed, nz)
    ccs = ns.Ccs(inst)
    pp = ns.PublicParams(args.n, seed=PP_SEED)
    sess = ns.FoldingSession(ccs, pp)
    z_dev = [ns.DeviceBuffer.from_array(z) for z in zs]
    ptrs = [b.data_ptr() for b in z_dev]
    ns.device_synchronize()
    setup_s = time.perf_counter() - t_setup

    if args.warmup > 0:
        sess.prove_device(ptrs[:args.warmup])
    # The commitment autotuner times its K-split candidates on the first commit of each
    # (m, witnesses) shape, as a long-running prover does on its first proofs: the warmup settled
    # the 12-child shape; settle the batched commitment of the timed run's steps here (zero
    # digits cost the GEMM what real ones do).
    zeros = [ns.DeviceBuffer.from_array(np.zeros(54 * args.n, np.int8)) for _ in range(args.steps)]
    pp.commit_device(zeros)
    del zeros
    ns.device_synchronize()
    t0 = time.perf_counter()
    proof = sess.prove_device(ptrs[:args.steps])
    ns.device_synchronize()
    t1 = time.perf_counter()
    elapsed = t1 - t0
    phases = sess.phase_stats()
    perms = sess.phase_perms()
    # kernel durations from a second, identical proof with per-launch start/stop events (kept out
    # of the timed run, whose launches carry no events)
    sess.set_timing(True)
    t2 = time.perf_counter()
    proof_timed = sess.prove_device(ptrs[:args.steps])
    ns.device_synchronize()
    elapsed_timed = time.perf_counter() - t2
    sess.set_timing(False)
    ms, counts = sess.kernel_stats()
    if proof_timed != proof:
        sys.exit("the event-timed proof differs from the timed run's")

    work = algorithmic_work(args.n, args.steps, args.steps)
    dom = max(range(len(KT)), key=lambda i: ms[i])
    cls = KT[dom]
    bound, unit, peak = kernel_bound(cls)
    avg_ms = ms[dom] / max(counts[dom], 1)
    per_launch = work[cls] / max(counts[dom], 1)
    achieved = (per_launch / (avg_ms * 1e-3)) / (1e9 if unit == "GB/s" else 1e12)
    # HBM bytes per launch of this kernel from the committed FETCH_SIZE / WRITE_SIZE summary
    # (copy-kernel calibrated), taken on the 2^16-row, 16-step shard; per-launch traffic of a
    # round kernel depends on n, not on the step count
    traffic = None
    if args.n == 1 << 16:
        from nightstream_amd.profiling import load_traffic
        names = KERNEL_NAMES[cls]
        if cls == "ajtai_gemm":
            # the commitment tuner times every tile shape on its first commit; the timed region
            # runs the shape it chose, which the committed kernel summary names ("... xN")
            import re
            try:
                tail = open(os.path.join(ROOT, "profiles", "kernel_stats_plain_shard_2e16.md")).read()
                chosen = tuple(re.findall(r"(k_ajtai_gemm\w*<?\d*>?) x\d+", tail))
                names = chosen or names
            except OSError:
                pass
        traffic = load_traffic(os.path.join(ROOT, "profiles", "traffic_plain_shard_2e16.md"), names)
    roofline = {
        "bound": bound, "achieved": round(achieved, 3), "peak": peak, "unit": unit,
        "frac": round(achieved / peak, 5), "traffic": traffic,
        "kernel": "+".join((KERNEL_NAMES[cls],) if isinstance(KERNEL_NAMES[cls], str) else KERNEL_NAMES[cls]),
        "launches": int(counts[dom]), "avg_ms": round(avg_ms, 5),
        "per_launch_algorithmic": per_launch,
        "kernel_ms": {KT[i]: round(ms[i], 4) for i in range(len(KT))},
    }

    # the same figures for every kernel class with measured time, largest summed time first (the
    # roofline object above is the first of them)
    by_kernel = []
    tot_ms = sum(ms)
    for i in sorted(range(len(KT)), key=lambda i: -ms[i]):
        if counts[i] == 0:
            continue
        b_i, u_i, p_i = kernel_bound(KT[i])
        avg_i = ms[i] / counts[i]
        ach_i = (work[KT[i]] / counts[i] / (avg_i * 1e-3)) / (1e9 if u_i == "GB/s" else 1e12)
        by_kernel.append({"kernel": KT[i], "bound": b_i, "unit": u_i, "achieved": round(ach_i, 1),
                          "frac": round(ach_i / p_i, 4), "avg_ms": round(avg_i, 5), "launches": int(counts[i]),
                          "share_of_kernel_time": round(ms[i] / tot_ms, 3)})

    cpu = None
    parity = None
    if args.cpu_steps > 0:
        cpu, cpu_proof = cpu_baseline(inst, zs, args.cpu_steps)
        # the GPU proves the same first steps as a shard of their own (outside the timed region)
        gpu_proof = ns.FoldingSession(ccs, pp).prove_device(ptrs[:args.cpu_steps])
        parity = {"steps": args.cpu_steps, "proof_bytes": len(cpu_proof), "bit_identical": gpu_proof == cpu_proof}

    value = args.steps / elapsed
    line = {
        "metric": "fold-steps/s at 2^16-row CCS chunk vs CPU; p50/p99 job latency under load",
        "value": round(value, 4),
        "unit": "fold-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u64 (Goldilocks F_q / F_q^2; int8 MFMA with int32 accumulation for exact contractions)",
        "data": "synthetic",
        "config": {
            "workload": f"plain shard: R1CS n = m = 2^{args.n.bit_length() - 1}, {args.steps} fold steps with "
                        "distinct witnesses, default Goldilocks params (D=54, kappa=16, b=2, k_rho=12), CCS-only",
            "n": args.n, "m": args.n, "fold_steps": args.steps, "m_in": 1,
            "proof_bytes": len(proof), "setup_s": round(setup_s, 2),
        },
        "roofline": roofline,
        "roofline_by_kernel": by_kernel,
        "cpu_baseline": cpu,
        "parity_vs_cpu": parity,
        "gpu_kernel_ms_per_step": round(sum(ms) / args.steps, 4),
        "ms_per_step_with_kernel_events": round(1e3 * elapsed_timed / args.steps, 4),
        "host_phases_s": phases,
        "host_phase_perms": perms,
    }
    if args.mix:
        del sess, z_dev
        line["mix"] = mix_measurement(args.workers, args.seed)
    print(json.dumps(line), flush=True)
    if parity is not None and not parity["bit_identical"]:
        sys.exit("GPU proof differs from the CPU restatement's on the sampled steps")
    if args.mix and not (line["mix"]["proofs_bit_identical"] and line["mix"]["serve"]["served_equal_plain"]):
        sys.exit("pooled or served proofs differ from the plain path")


def _stderr(msg):
    print(msg, file=sys.stderr, flush=True)


def batch_specs(rng, shards: int, lo: int, hi: int, smin: int, smax: int, memory: bool, variants: int = 4):
    """(n, steps, seed, memory) specs: rows 2^lo..2^hi log-uniform, smin..smax fold steps, up to
    `variants` distinct instances per size; with `memory`, every other shard is a Twist/Shout
    shard (a RISC-V-like trace of n / 128 cycles per step)."""
    shape = [(int(rng.integers(lo, hi + 1)), int(rng.integers(smin, smax + 1))) for _ in range(shards)]
    return [(1 << e, steps, 1000 + e * 16 + (i % variants), bool(memory and i % 2 == 0))
            for i, (e, steps) in enumerate(shape)]


def run_batch(specs, workers: int, procs: int = 1, hw_queues: int = 1):
    """The plain path (one shard at a time in one prover context) against `workers` concurrent
    contexts (in `procs` prover processes); both keep contexts warmed before their timed run.
    Returns (record, shards, proofs)."""
    import nightstream_amd as ns
    from nightstream_amd.batch import ProcessProver, ProverPool, ShardFactory, build_shards, prove_batched, \
        prove_sequential

    # longest estimated time first (a Twist/Shout step ~0.2 s of host transcript at any size)
    order = sorted(range(len(specs)), key=lambda i: -specs[i][1] * (0.2 if specs[i][3] else
                                                                    0.012 * specs[i][0] / 65536 + 0.002))
    t0 = time.perf_counter()
    pp = None
    if procs > 1:  # started before this process touches the GPU; each builds and warms its own
        pp = ProcessProver(PP_SEED, specs, procs, max(1, workers // procs), hw_queues)
    t_procs = time.perf_counter() - t0
    t0 = time.perf_counter()
    shards = build_shards(ShardFactory(PP_SEED), specs)
    ns.device_synchronize()
    setup = time.perf_counter() - t0
    total_steps = sum(s.steps for s in shards)
    mem_steps = sum(s.steps for s in shards if s.mem is not None)
    _stderr(f"batch: {len(shards)} shards, {total_steps} fold steps ({mem_steps} with Twist/Shout memory), "
            f"setup {setup:.1f} s, {procs} prover processes ready in {t_procs:.1f} s")
    pool1 = ProverPool(1)
    t0 = time.perf_counter()
    pool1.warm(shards)
    pooln = None
    if pp is None:
        pooln = ProverPool(workers)
        pooln.warm(shards)
    t_warm = time.perf_counter() - t0
    seq, t_seq = prove_sequential(shards, pool1)
    _stderr(f"batch: plain path {t_seq:.2f} s")
    hub0 = ns.shared_round_stats()
    if pp is None:
        bat, t_bat = prove_batched(shards, workers, pooln)
        created = pooln.created - workers
    else:
        bat, t_bat = pp.prove_all(order)
        created = 0
        pp.close()
    _stderr(f"batch: {workers} contexts in {procs} processes {t_bat:.2f} s")
    rec = {
        "shards": len(shards), "fold_steps": total_steps, "fold_steps_with_memory": mem_steps,
        "workers": workers, "processes": procs,
        "plain_fold_steps_per_s": round(total_steps / t_seq, 3),
        "batched_fold_steps_per_s": round(total_steps / t_bat, 3),
        "speedup_vs_plain": round(t_seq / t_bat, 3),
        "seconds": {"plain": round(t_seq, 3), "batched": round(t_bat, 3), "setup": round(setup, 1),
                    "context_warmup": round(t_warm, 2), "process_start": round(t_procs, 1)},
        "contexts_created_in_timed_runs": pool1.created - 1 + created,
        "proofs_bit_identical": all(a == b for a, b in zip(seq, bat)),
    }
    hub1 = ns.shared_round_stats()
    if pp is None and hub1[1] > hub0[1]:
        # fused Pi_CCS rounds of concurrent contexts issued together (this process only)
        rec["shared_round_launches"] = {"rounds": hub1[1] - hub0[1], "launches": hub1[0] - hub0[0],
                                        "rounds_per_launch": round((hub1[1] - hub0[1]) / max(1, hub1[0] - hub0[0]), 2)}
    return rec, shards, seq, (pooln or pool1)


def bench_batch(args):
    """128 independent shards (rows 2^10..2^16 log-uniform, 4-16 fold steps, a distinct witness
    per step; every other one a Twist/Shout shard whose steps carry Route A memory over the shared
    CPU bus): the plain path (one shard at a time in one prover context) against --workers
    concurrent contexts spread over --procs prover processes; every proof of the batched run must
    equal the plain run's."""
    rng = np.random.default_rng(args.seed)
    specs = batch_specs(rng, args.shards, 10, 16, 4, 16, bool(args.batch_memory))
    rec, _, _, _ = run_batch(specs, args.workers, max(1, args.procs), args.proc_hw_queues)
    line = {
        "metric": "fold-steps/s, 128-shard batch (2^10..2^16 rows, 4-16 steps, half with Twist/Shout memory)",
        "value": rec["batched_fold_steps_per_s"], "unit": "fold-steps/s", "n_gpus": 1,
        "higher_is_better": True, "vs_baseline": None, "dtype": "u64", "data": "synthetic",
        "config": {"workload": "batch", "shards": rec["shards"], "fold_steps": rec["fold_steps"],
                   "fold_steps_with_memory": rec["fold_steps_with_memory"], "workers": args.workers,
                   "processes": rec["processes"],
                   "hw_queues_per_process": args.proc_hw_queues if rec["processes"] > 1 else None,
                   "specs": [[n.bit_length() - 1, st, int(m)] for n, st, _, m in specs]},
        **{k: v for k, v in rec.items() if k not in ("shards", "fold_steps", "fold_steps_with_memory", "workers",
                                                     "processes")},
    }
    print(json.dumps(line), flush=True)
    if not rec["proofs_bit_identical"]:
        sys.exit("batched proofs differ from the plain path")


def serve_rate(jobs, rate, duration, workers, pool, pp_res, seed):
    """One open-loop Poisson run; returns its record and results."""
    from nightstream_amd.batch import serve
    res = serve(jobs, rate, duration, workers, 240e9, seed=seed, pp_resident=pp_res, pool=pool)
    if not res:
        return None, res
    lat = np.array([r.latency for r in res])
    span = max(r.finish for r in res) - min(r.arrival for r in res)
    steps = sum(r.steps for r in res)
    return {"jobs": len(res), "fold_steps_per_s": round(steps / span, 3),
            "p50_s": round(float(np.percentile(lat, 50)), 3),
            "p99_s": round(float(np.percentile(lat, 99)), 3)}, res


def bench_serve(args):
    """Open-loop Poisson arrivals; jobs of 2^12..2^18 rows (log-uniform) with 1-32 fold steps
    (a distinct witness per step), --serve-memory-frac of them Twist/Shout shards; admission by
    free worker and by estimated HBM footprint under a 240 GB cap.  For each rate: fold-steps/s
    and p50/p99 latency, for the pooled server and for the plain single-context server; the sweep
    stops once the pooled p99 exceeds 5 s."""
    import nightstream_amd as ns
    from nightstream_amd.batch import ProverPool, ShardFactory, pp_bytes

    rng = np.random.default_rng(args.seed)
    fac = ShardFactory(PP_SEED)
    t0 = time.perf_counter()
    base = {}
    for e in range(12, 19):
        for v in range(2):
            base[(e, v)] = fac.make(1 << e, 32, 2000 + e * 4 + v)
        if args.serve_memory_frac > 0:
            base[(e, "mem")] = fac.make(1 << e, 32, 2000 + e * 4 + 2, memory=True)
    ns.device_synchronize()
    jobs = []
    for i in range(4096):
        e = int(rng.integers(12, 19))
        b = base[(e, "mem")] if rng.random() < args.serve_memory_frac else base[(e, i % 2)]
        jobs.append(b.with_steps(int(rng.integers(1, 33))))
    pp_res = float(sum(pp_bytes(1 << e) for e in range(12, 19)))
    _stderr(f"serve: pool of {len(base)} instances ready in {time.perf_counter() - t0:.1f} s")
    solo = {}
    rates = [float(r) for r in args.rates.split(",")] if args.rates else [5.0, 10.0, 20.0, 40.0, 80.0]
    out = []
    # each server keeps its contexts across rates (as a running server does), created and sized
    # for the largest job before the first rate
    pools = {"pooled": ProverPool(args.workers), "plain": ProverPool(1)}
    for p in pools.values():
        p.warm(jobs)
    plain_saturated = False  # once the single-context server's p99 passes 10 s its queue only grows
    for rate in rates:
        rec = {"rate_jobs_per_s": rate}
        for label, workers in (("pooled", args.workers), ("plain", 1)):
            if label == "plain" and plain_saturated:
                rec["plain"] = "saturated at a lower rate"
                continue
            r, res = serve_rate(jobs, rate, args.rate_duration, workers, pools[label], pp_res, int(rate * 1000))
            if r is None:
                continue
            rec[label] = r
            if label == "pooled":
                for x in res[:8]:
                    solo.setdefault(x.job, x.proof)
            _stderr(f"serve: rate {rate} {label}: {r}")
            if label == "plain" and r["p99_s"] > 10.0:
                plain_saturated = True
            if label == "pooled" and r["p99_s"] > 5.0:
                break
        out.append(rec)
        if rec.get("pooled", {}).get("p99_s", 0) > 5.0:
            break
    # each mixed result equals the same job proved alone
    alone = ProverPool(1)
    same = all(alone.prove(jobs[j]) == p for j, p in solo.items())
    line = {
        "metric": "served fold-steps/s and p50/p99 job latency (Poisson arrivals, 2^12..2^18 rows, 1-32 steps, "
                  f"{args.serve_memory_frac:.0%} with Twist/Shout memory)",
        "value": max((r["pooled"]["fold_steps_per_s"] for r in out if "pooled" in r), default=None),
        "unit": "fold-steps/s", "n_gpus": 1, "higher_is_better": True, "vs_baseline": None, "dtype": "u64",
        "data": "synthetic",
        "config": {"workload": "serve", "workers": args.workers, "rate_duration_s": args.rate_duration,
                   "hbm_cap_bytes": 240e9, "memory_frac": args.serve_memory_frac},
        "rates": out, "mixed_equals_alone": same,
    }
    print(json.dumps(line), flush=True)


def mix_measurement(workers: int, seed: int):
    """Short in-process evidence for the batch and serve claims beside the shard line: 48 shards
    (2^10..2^13 rows, 3-8 fold steps, every other one a Twist/Shout shard; three per pooled
    context, so that the batched time is not one shard's serial chain) proved on the plain path
    and through `workers` pooled contexts (proofs compared), then ~8 s of Poisson arrivals over the
    same shards at about 60% of the pooled throughput, with p50/p99 job latency."""
    from nightstream_amd.batch import pp_bytes
    rng = np.random.default_rng(seed)
    specs = batch_specs(rng, 48, 10, 13, 3, 8, True, variants=2)
    rec, shards, seq, pool = run_batch(specs, workers)
    mean_steps = rec["fold_steps"] / rec["shards"]
    rate = round(0.6 * rec["batched_fold_steps_per_s"] / mean_steps, 2)
    pp_res = float(sum(pp_bytes(1 << e) for e in range(10, 14)))
    srv, res = serve_rate(shards, rate, 8.0, workers, pool, pp_res, seed)
    same = all(r.proof == seq[r.job % len(shards)] for r in res)
    rec["serve"] = {"rate_jobs_per_s": rate, "duration_s": 8.0, **(srv or {}), "served_equal_plain": same}
    rec["specs"] = [[n.bit_length() - 1, st, int(m)] for n, st, _, m in specs]
    return rec


Q_FIELD = 0xFFFFFFFF00000001


def bench_twist(args):
    """Memory sidecar of a RISC-V-like trace, Twist/Shout round oracles only: 2^12-cycle chunks; every oracle
    of a chunk is constructed, the oracles are grouped by number of rounds (12 time-lane rounds,
    5 or 16 address-lane rounds) and each group is driven through all its rounds with shared
    seeded challenges (no transcript), on the GPU with the batched driver, timed per chunk.  The
    same chunks are also timed one oracle at a time on the GPU (and built on one thread); the CPU
    restatement runs chunk 0 as the baseline, and chunk 0's values must be identical on all three.
    With the batched driver a chunk's oracles are built on --build-threads host threads."""
    import nightstream_amd as ns
    from nightstream_amd import twist
    from nightstream_amd.sidecar import chunk_lanes, chunk_oracles, drive_group
    from oracle import pyoracle
    rng_ch = np.random.default_rng(args.seed)
    chunks = [chunk_lanes(c) for c in range(args.steps)]
    chals = [[(int(a) * 2 % Q_FIELD, int(b) * 2 % Q_FIELD) for a, b in rng_ch.integers(0, 1 << 63, size=(40, 2))]
             for _ in range(64)]

    def run(cls, c, batch=None, pool=None):
        twists, shouts = chunks[c]
        rng = np.random.default_rng(1000 + c)
        t0 = time.perf_counter()
        makers = chunk_oracles(cls, twists, shouts, rng)
        if pool is not None:  # the constructors release the GIL; each oracle owns its stream
            made = list(pool.map(lambda lm: (lm[0], lm[1]()), makers))
        else:
            made = [(label, make()) for label, make in makers]
        per = {"construct": time.perf_counter() - t0}
        groups = {}
        for i, (_, o) in enumerate(made):
            groups.setdefault(o.num_rounds(), []).append(i)
        vals = [[] for _ in made]
        for nr, idxs in groups.items():
            t1 = time.perf_counter()
            for i, v in zip(idxs, drive_group([made[i][1] for i in idxs], chals[nr], batch)):
                vals[i] = v
            per[f"{nr} rounds x {len(idxs)} oracles"] = time.perf_counter() - t1
        return vals, per, len(made)

    def timed(batch, pool=None):
        for c in range(min(args.warmup, len(chunks))):
            run(twist.TsOracle, c, batch, pool)
        ns.device_synchronize()
        t0 = time.perf_counter()
        per_all, vals0, n_or = {}, None, 0
        for c in range(len(chunks)):
            v, per, n_or = run(twist.TsOracle, c, batch, pool)
            if c == 0:
                vals0 = v
            for k, x in per.items():
                per_all[k] = per_all.get(k, 0.0) + x
        ns.device_synchronize()
        return time.perf_counter() - t0, per_all, vals0, n_or

    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=args.build_threads) as pool:
        t_gpu, per_all, gpu_vals0, n_or = timed(twist, pool)
    t_one, _, one_vals0, _ = timed(None)
    pyoracle.build()
    t1 = time.perf_counter()
    cpu_vals0, _, _ = run(pyoracle.TsOracle, 0)
    t_cpu = time.perf_counter() - t1
    identical = cpu_vals0 == gpu_vals0 == one_vals0
    line = {
        "metric": "Twist/Shout round oracles of a RISC-V-like trace chunk (2^12 cycles), chunks/s",
        "value": round(len(chunks) / t_gpu, 3), "unit": "chunks/s", "n_gpus": 1, "steps": len(chunks),
        "higher_is_better": True, "vs_baseline": None, "dtype": "u64 (Goldilocks F_q^2)", "data": "synthetic",
        "config": {"workload": "twist: 2 register-file Twists (32 words), RAM Twist (2^16 words, 35%, Zipf 1.1, "
                               "40% writes), 3 Shout lanes (8x8-bit ADD/XOR/AND); 2^12-cycle chunks; oracles "
                               "grouped by number of rounds, each group a batched sum-check",
                   "oracles_per_chunk": n_or, "chunks": len(chunks)},
        "ms_per_chunk": round(t_gpu / len(chunks) * 1e3, 3),
        "ms_per_chunk_one_oracle_at_a_time": round(t_one / len(chunks) * 1e3, 3),
        "construct_threads": args.build_threads,
        "ms_per_phase_per_chunk": {k: round(v / len(chunks) * 1e3, 3) for k, v in per_all.items()},
        "cpu_baseline": {"value": round(1.0 / t_cpu, 4), "unit": "chunks/s", "cores": 1, "kind": "port",
                         "sample": "CPU restatement (oracle/twist.c) of chunk 0, every oracle, single thread"},
        "values_bit_identical_chunk0": identical,
    }
    print(json.dumps(line), flush=True)
    if not identical:
        sys.exit("GPU and CPU Twist/Shout values differ")


def bench_route_a(args):
    """Shard with Twist/Shout memory over the shared CPU bus: a RISC-V-like trace in
    2^log_chunk-cycle chunks, one fold step per chunk (nightstream_amd/bus_workload.py: 2 register
    Twists, the RAM Twist, 3 Shout lookup lanes, the bus in an R1CS of n = m = 128 * chunk).  The
    timed region proves --steps steps with the witnesses resident in HBM; the proof carries the
    memory sidecar and both obligation lanes.  The CPU restatement proves the first --cpu-steps
    steps as the baseline and its proof of them must equal the GPU's."""
    import nightstream_amd as ns
    from nightstream_amd.bus_workload import gen_bus_shard
    if ns.device_count() < 1:
        sys.exit("no GPU visible")
    t_setup = time.perf_counter()
    nz = max(args.steps, args.warmup, args.cpu_steps, 1)
    inst, zs, mems, lay = gen_bus_shard(args.log_chunk, nz, args.seed)
    ccs = ns.Ccs(inst)
    pp = ns.PublicParams(inst.m, seed=PP_SEED)
    sess = ns.FoldingSession(ccs, pp)
    bufs = [ns.DeviceBuffer.from_array(z) for z in zs]
    ptrs = [b.data_ptr() for b in bufs]
    ns.device_synchronize()
    setup_s = time.perf_counter() - t_setup
    if args.warmup > 0:
        sess.prove_device(ptrs[:args.warmup], mem=mems[:args.warmup])
    zeros = [ns.DeviceBuffer.from_array(np.zeros(54 * inst.m, np.int8)) for _ in range(args.steps)]
    pp.commit_device(zeros)
    del zeros
    ns.device_synchronize()
    sess.set_timing(True)
    t0 = time.perf_counter()
    proof = sess.prove_device(ptrs[:args.steps], mem=mems[:args.steps])
    ns.device_synchronize()
    elapsed = time.perf_counter() - t0
    sess.set_timing(False)
    ms, counts = sess.kernel_stats()
    phases = sess.phase_stats()
    perms = sess.phase_perms()
    cpu, parity = None, None
    if args.cpu_steps > 0:
        from oracle import pyoracle
        pyoracle.build()
        prover = pyoracle.Prover(inst, PP_SEED)
        t1 = time.perf_counter()
        cpu_proof, _ = prover.prove(zs[:args.cpu_steps], mem=mems[:args.cpu_steps])
        wall = time.perf_counter() - t1
        secs = prover.step_seconds
        idx = 1 if args.cpu_steps >= 2 else 0
        cpu = {"value": 1.0 / secs[idx], "unit": "fold-steps/s", "cores": pyoracle.num_threads(), "kind": "port",
               "sample": f"CPU restatement (oracle/) of the first {args.cpu_steps} steps: step {idx} took "
                         f"{secs[idx]:.2f} s ({wall:.1f} s wall with the MCS commits)"}
        gpu_proof = ns.FoldingSession(ccs, pp).prove_device(ptrs[:args.cpu_steps], mem=mems[:args.cpu_steps])
        parity = {"steps": args.cpu_steps, "proof_bytes": len(cpu_proof), "bit_identical": gpu_proof == cpu_proof}
    line = {
        "metric": "fold-steps/s of a Twist/Shout shard (RISC-V-like trace, 2^12-cycle chunks)",
        "value": round(args.steps / elapsed, 4), "unit": "fold-steps/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(1e3 * elapsed / args.steps, 3), "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "u64 (Goldilocks F_q / F_q^2; int8 MFMA)",
        "data": "synthetic",
        "config": {"workload": f"route_a: 2^{args.log_chunk}-cycle chunks, {args.steps} fold steps; 2 register "
                               "Twists (32 words), RAM Twist (2^16 words, 35% accesses, Zipf 1.1, 40% writes), 3 Shout "
                               "lanes (8x8-bit ADD/XOR/AND); shared bus of 121 columns in an R1CS of "
                               f"n = m = {inst.n}",
                   "n": inst.n, "bus_cols": lay.bus_cols, "chunk": lay.chunk, "proof_bytes": len(proof),
                   "setup_s": round(setup_s, 2)},
        "kernel_ms": {KT[i]: round(ms[i], 4) for i in range(len(KT))},
        "gpu_kernel_ms_per_step": round(sum(ms) / args.steps, 4),
        "host_phases_s": phases, "host_phase_perms": perms,
        "cpu_baseline": cpu, "parity_vs_cpu": parity,
    }
    print(json.dumps(line), flush=True)
    if parity is not None and not parity["bit_identical"]:
        sys.exit("GPU proof differs from the CPU restatement's on the sampled steps")


AJTAI_MACS_PER_COLUMN = 16 * 54 * 54 * 8  # kappa * D^2 * 8 byte limbs: int8 MACs per column per witness


def bench_ajtai(args):
    """Ajtai commitment sweep: 64 witnesses of m = 2^16 .. 2^20 columns (balanced base-2 digits
    uniform in {-1, 0, 1}, D = 54 rows) committed in one batched launch against PP sampled from
    the seed on the GPU, with the digit matrices already resident in HBM; timed per m.  Parity
    at every m: a sparse witness (digits in 64 random columns) against the oracle on those PP
    columns (a commitment is a sum over columns); at 2^16 also a dense witness of the batch
    against the oracle, whose time is the CPU baseline."""
    import nightstream_amd as ns
    from oracle import pyoracle
    pyoracle.build()
    nw, stride = 64, 4096
    logms = list(range(16, args.ajtai_max_log + 1))
    seed = bytes([7] * 32)
    rng = np.random.default_rng(args.seed)
    maxm = 1 << max(logms)
    # witness w is the window [w*stride, w*stride + 54 m) of one uniform pool
    pool = rng.integers(-1, 2, size=54 * maxm + stride * (nw - 1), dtype=np.int8)
    per_m, ok, t_cpu = {}, True, None
    for lm in logms:
        m = 1 << lm
        t0 = time.perf_counter()
        pp = ns.PublicParams(m, seed=seed, keep_coeffs=True)
        ns.device_synchronize()
        t_setup = time.perf_counter() - t0
        bufs = [ns.DeviceBuffer.from_array(np.ascontiguousarray(pool[w * stride: w * stride + 54 * m]))
                for w in range(nw)]
        first = pp.commit_device(bufs)
        reps = 3 if lm >= 19 else 6
        t0 = time.perf_counter()
        for _ in range(reps):
            c = pp.commit_device(bufs)
        t = (time.perf_counter() - t0) / reps
        repeat_ok = np.array_equal(c, first)
        # each tile shape alone (its own K-split tuning), then back to the timed choice
        shape_ms = {}
        for v, name in ((1, "16x16x64"), (2, "32x32x32"), (3, "32x32x32 packed digits")):
            ns.ajtai_force_variant(v)
            cv = pp.commit_device(bufs)
            t0 = time.perf_counter()
            for _ in range(reps):
                cv = pp.commit_device(bufs)
            shape_ms[name] = round((time.perf_counter() - t0) / reps * 1e3, 3)
            repeat_ok &= np.array_equal(cv, first)
        ns.ajtai_force_variant(0)
        cols = np.sort(rng.choice(m, size=64, replace=False))
        zsub = rng.integers(-1, 2, size=(54, 64), dtype=np.int8)
        zs = np.zeros((54, m), np.int8)
        zs[:, cols] = zsub
        got = pp.commit_device([ns.DeviceBuffer.from_array(zs.reshape(-1))])[0]
        ppc = pp.coeffs().reshape(16, m, 54)
        want = pyoracle.ajtai_commit(np.ascontiguousarray(ppc[:, cols, :]).reshape(-1), 16, 64, zsub)
        sparse_ok = np.array_equal(got, want)
        dense_ok = None
        if lm == logms[0]:
            t1 = time.perf_counter()
            want0 = pyoracle.ajtai_commit(ppc.reshape(-1), 16, m, pool[:54 * m].reshape(54, m))
            t_cpu = (time.perf_counter() - t1) / m  # seconds per column of one witness
            dense_ok = np.array_equal(first[0], want0)
        ok &= repeat_ok and sparse_ok and dense_ok is not False
        per_m[f"2^{lm}"] = {"ms_per_64": round(t * 1e3, 3), "witnesses_per_s": round(nw / t, 2),
                            "int8_TOPS": round(2 * AJTAI_MACS_PER_COLUMN * m * nw / t / 1e12, 1),
                            "tile_shape_ms_per_64": shape_ms,
                            "pp_setup_s": round(t_setup, 3), "sparse_parity": sparse_ok,
                            **({"dense_parity": dense_ok} if dense_ok is not None else {})}
        print(f"ajtai m=2^{lm}: {per_m[f'2^{lm}']}", file=sys.stderr, flush=True)
        del bufs, pp, ppc
    top = f"2^{max(logms)}"
    tops = per_m[top]["int8_TOPS"]
    line = {
        "metric": f"Ajtai commitments of 64 witnesses sharing PP (2^16..2^20 sweep), witnesses/s at m = {top}",
        "value": per_m[top]["witnesses_per_s"], "unit": "witnesses/s", "n_gpus": 1, "steps": len(logms),
        "higher_is_better": True, "vs_baseline": None,
        "dtype": "int8 MFMA, int32 accumulation, exact Goldilocks reduction", "data": "synthetic",
        "config": {"workload": "ajtai sweep: 64 witnesses, m = 2^16..2^20, balanced base-2 digits uniform in "
                               "{-1,0,1}, kappa = 16, D = 54, PP from a 32-byte seed", "witnesses": nw},
        "roofline": {"bound": "mfma", "achieved": tops, "peak": 5000.0, "unit": "TOP/s", "frac": round(tops / 5000.0, 4),
                     "traffic": None, "note": "whole commit call (GEMM, split sums, epilogue, allocations, "
                                              "commitment copy) at the top m"},
        "per_m": per_m,
        "cpu_baseline": {"value": round(1.0 / (t_cpu * (1 << max(logms))), 6), "unit": "witnesses/s", "cores": 1,
                         "kind": "port", "sample": f"oracle commit of one dense witness at m = 2^{logms[0]} "
                                                   f"({t_cpu * (1 << logms[0]):.2f} s), scaled to {top} columns "
                                                   "(the cost is linear in m)"},
        "bit_exact": ok,
    }
    print(json.dumps(line), flush=True)
    if not ok:
        sys.exit("Ajtai commitments differ from the oracle")


if __name__ == "__main__":
    main()
