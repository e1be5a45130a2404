"""Batched and served proving: many independent shards on one GPU.

A shard's fold steps are sequential (each step's accumulator feeds the next), and each step's
Fiat-Shamir transcript is host work of roughly constant size (about 17.8k Poseidon2
permutations for a CCS-only step, independent of n; a Twist/Shout step binds every ME's 121 bus
openings too, about 20x more).  Throughput across shards therefore comes from running many
sessions at once: a running job holds one prover context (its own HIP stream and device buffers,
kept in a pool and reused from job to job through `rebind`), drives its shard's transcript on its
own core, and its kernels overlap with the other jobs' on the GPU.  ctypes releases the GIL for the
whole `prove` call, so the Python layer only dispatches.

  prove_sequential(shards)          the plain path: one shard at a time, one prover context
  prove_batched(shards, workers)    the same shards through a pool of `workers` contexts
  ProverPool(workers).warm(shards)  contexts created and sized up front, reusable across calls
  serve(jobs, rate, ...)            open-loop Poisson arrivals with continuous admission by slot
                                    and by estimated HBM footprint; per-job latency
Every proof is the same byte string whichever path produced it (each shard has its own
transcript and the kernels are deterministic).
"""
from __future__ import annotations

import contextlib
import dataclasses
import queue
import threading
import time

import numpy as np

from .workload import Q, R1CS, rand_balanced_z, rand_field

D = 54


def gen_r1cs_fast(n: int, seed: int):
    """The R1CS generator's distribution (A, B: 3 distinct uniform columns per row with uniform
    nonzero coefficients; C: one column j != 0 per row with coefficient (Az)_i (Bz)_i / z_j;
    z[0] = 1, other entries uniform in |z| <= 2^54 - 1), vectorised for large batches."""
    rng = np.random.default_rng(seed)
    m = n
    z = rand_balanced_z(rng, m)
    zo = z.astype(object)

    def rows3():
        cols = rng.integers(0, m, size=(n, 3))
        while True:
            dup = (cols[:, 0] == cols[:, 1]) | (cols[:, 0] == cols[:, 2]) | (cols[:, 1] == cols[:, 2])
            if not dup.any():
                break
            cols[dup] = rng.integers(0, m, size=(int(dup.sum()), 3))
        cols.sort(axis=1)
        vals = rand_field(rng, 3 * n).reshape(n, 3)
        return cols.astype(np.uint32), vals

    a_cols, a_vals = rows3()
    b_cols, b_vals = rows3()
    az = (a_vals.astype(object) * zo[a_cols.astype(np.int64)]).sum(axis=1) % Q
    bz = (b_vals.astype(object) * zo[b_cols.astype(np.int64)]).sum(axis=1) % Q
    jj = rng.integers(1, m, size=n)
    zero = z[jj] == 0
    while zero.any():
        jj[zero] = 1 + (jj[zero] % (m - 1))
        zero = z[jj] == 0
    inv = np.array([pow(int(v), Q - 2, Q) for v in z[jj]], dtype=object)
    c_vals = np.array((az * bz % Q) * inv % Q, dtype=np.uint64)
    rp3 = np.arange(0, 3 * n + 1, 3, dtype=np.uint64)
    rp1 = np.arange(0, n + 1, dtype=np.uint64)
    a = (rp3, a_cols.reshape(-1).copy(), a_vals.reshape(-1).copy())
    b = (rp3, b_cols.reshape(-1).copy(), b_vals.reshape(-1).copy())
    c = (rp1, jj.astype(np.uint32), c_vals)
    return R1CS(n, m, a, b, c, z), z


def footprint_bytes(n: int, k: int = 13) -> int:
    """HBM a proving job needs beyond the shared PP: digit planes of k witnesses and 2 x 16 child
    planes, the round-1 NC tables (double-buffered), Z_mix and the scalar tables."""
    m = n
    return int(k * D * m + 2 * 16 * D * m + 2 * (k * D + 6) * 16 * (n // 2) + 4 * D * m + 8 * 16 * n)


def pp_bytes(m: int) -> int:
    return 6912 * m


def mem_footprint_bytes(n: int) -> int:
    """HBM that Twist/Shout memory adds to a job of n = 128 c rows (c cycles per step): the val
    lane's 16 digit planes and its Z_mix, the decoded bus region and the bus weights, and the
    memory oracles (at most 4 KiB per cycle) with the three 2^16-entry lookup tables."""
    c = n // 128
    return int(16 * D * n + 4 * D * n + 8 * n + 16 * 121 * c + 4096 * c + 3 * 16 * (1 << 16))


@dataclasses.dataclass
class Shard:
    """One shard: `steps` fold steps, each over its own witness (z = [x | w], u64[m]) already in
    HBM; with `mem`, a Twist/Shout shard whose per-step memory instances ride the shared CPU bus
    of its witnesses (nightstream_amd/bus_workload.py).  z_devs and mem may hold more steps than
    `steps` (shards of one trace share them); the first `steps` are proved."""
    n: int
    steps: int
    seed: int
    ccs: object = None
    pp: object = None
    z_devs: list = None  # DeviceBuffer per step
    mem: object = None  # StepMemArray of the per-step memory instances, or None

    def z_ptrs(self):
        return [b.data_ptr() for b in self.z_devs[:self.steps]]

    def with_steps(self, steps: int) -> "Shard":
        if steps > len(self.z_devs):
            raise ValueError("shard has fewer witnesses than steps")
        return Shard(self.n, steps, self.seed, self.ccs, self.pp, self.z_devs, self.mem)


class ShardFactory:
    """Builds shards on the device, sharing one Ajtai PP per width m."""

    def __init__(self, pp_seed: bytes):
        import nightstream_amd as ns
        self.ns = ns
        self.pp_seed = pp_seed
        self.pps = {}

    def pp(self, m):
        if m not in self.pps:
            self.pps[m] = self.ns.PublicParams(m, seed=self.pp_seed)
        return self.pps[m]

    @staticmethod
    def instance(n: int, steps: int, seed: int, memory: bool = False):
        """The host instance of a shard: (R1CS, [z per step], [StepMem per step] or None).  Without
        memory, an n-row R1CS whose steps differ in their free witness columns; with memory, a
        RISC-V-like trace of n / 128 cycles per step over Twist/Shout memory (n >= 2^8)."""
        if memory:
            from .bus_workload import gen_bus_shard
            e = n.bit_length() - 1
            if n != 1 << e or e < 8:
                raise ValueError("a Twist/Shout shard needs n = 128 * 2^k >= 256 rows")
            inst, zs, mems, _ = gen_bus_shard(e - 7, steps, seed)
            return inst, zs, mems
        from .workload import gen_r1cs_steps
        inst, zs = gen_r1cs_steps(n, seed, steps)
        return inst, zs, None

    def make(self, n: int, steps: int, seed: int, memory: bool = False) -> Shard:
        inst, zs, mems = self.instance(n, steps, seed, memory)
        sh = Shard(n, steps, seed)
        sh.ccs = self.ns.Ccs(inst)
        sh.pp = self.pp(n)
        sh.z_devs = [self.ns.DeviceBuffer.from_array(z) for z in zs]
        if mems is not None:
            sh.mem = self.ns.StepMemArray(mems)
        return sh


class ProverPool:
    """Up to `workers` prover contexts kept across jobs and calls: a job takes an idle context (or
    creates one while fewer than `workers` exist) and returns it when done, so a warmed pool proves
    without creating HIP streams or growing device buffers (both serialize in the HIP runtime)."""

    def __init__(self, workers: int):
        self.workers = workers
        self._idle = []
        self._lock = threading.Lock()
        self.created = 0

    def _acquire(self, shard: Shard):
        import nightstream_amd as ns
        with self._lock:
            s = self._idle.pop() if self._idle else None
            if s is None:
                self.created += 1
        if s is None:
            return ns.FoldingSession(shard.ccs, shard.pp)
        s.rebind(shard.ccs, shard.pp)
        return s

    def prove(self, shard: Shard) -> bytes:
        s = self._acquire(shard)
        try:
            return s.prove_device(shard.z_ptrs(), mem=shard.mem)
        finally:
            with self._lock:
                self._idle.append(s)

    def warm(self, shards):
        """Create every context and grow its buffers for the largest shapes in `shards`: the most
        rows and, independently, the most steps (the batched step commitment's workspace grows
        with the step count, not with the rows), and the largest Twist/Shout shard; proved
        `workers` times at once so that each proof takes its own context."""
        big = max(shards, key=lambda sh: (sh.n, sh.steps))
        steps = max(sh.steps for sh in shards)
        by_steps = max(shards, key=lambda sh: (len(sh.z_devs), sh.n))
        warm = [big.with_steps(min(len(big.z_devs), steps))]
        if len(by_steps.z_devs) >= steps and by_steps is not big:
            warm.append(by_steps.with_steps(steps))
        mem = [sh for sh in shards if sh.mem is not None]
        if mem:
            bm = max(mem, key=lambda sh: (sh.n, len(sh.z_devs)))
            warm.append(bm.with_steps(len(bm.z_devs)))
        sessions = [self._acquire(warm[0]) for _ in range(self.workers - len(self._idle))]
        with self._lock:
            sessions += self._idle
            self._idle = []

        def run(s):
            for w in warm:
                s.rebind(w.ccs, w.pp)
                s.prove_device(w.z_ptrs(), mem=w.mem)

        threads = [threading.Thread(target=run, args=(s,)) for s in sessions]
        for th in threads:
            th.start()
        for th in threads:
            th.join()
        with self._lock:
            self._idle.extend(sessions)


def prove_sequential(shards, pool: ProverPool | None = None):
    """The plain per-shard path: one context, one shard at a time."""
    pool = pool or ProverPool(1)
    out = []
    t0 = time.perf_counter()
    for sh in shards:
        out.append(pool.prove(sh))
    return out, time.perf_counter() - t0


# Pools can share fused-round launches among their concurrent contexts (proofs are the same
# either way).  Off by default: on the MI355X the rounds of 16 contexts coincide rarely (1.02
# rounds per shared launch on a 64-shard CCS-only batch), and issuing them on one shared stream
# gives up the overlap of the contexts' own streams (1,026 against 1,285 fold-steps/s).
SHARE_ROUNDS = False


@contextlib.contextmanager
def shared_rounds(on: bool):
    """Shared launches for the fused rounds of the contexts proving at once while the block runs
    (the plain path keeps one launch pair per round)."""
    import nightstream_amd as ns
    ns.set_shared_rounds(on and SHARE_ROUNDS)
    try:
        yield
    finally:
        ns.set_shared_rounds(False)


def est_seconds(shard) -> float:
    """Rough proving time of a shard for longest-first scheduling: a Twist/Shout step costs
    ~0.2 s of host transcript whatever its size (every ME binds its bus openings into Pi_RLC); a
    CCS-only step ~12 ms at 2^16 rows on the MI355X, scaling with the rows, plus ~2 ms of fixed
    host work."""
    per = 0.2 if shard.mem is not None else 0.012 * shard.n / 65536 + 0.002
    return shard.steps * per


def prove_batched(shards, workers: int, pool: ProverPool | None = None):
    """All shards through `workers` concurrent contexts (longest estimated time first, so the tail
    is short)."""
    import concurrent.futures as cf
    pool = pool or ProverPool(workers)
    order = sorted(range(len(shards)), key=lambda i: -est_seconds(shards[i]))
    out = [None] * len(shards)
    t0 = time.perf_counter()
    with shared_rounds(workers > 1), cf.ThreadPoolExecutor(max_workers=workers) as ex:
        futs = {ex.submit(pool.prove, shards[i]): i for i in order}
        for f in cf.as_completed(futs):
            out[futs[f]] = f.result()
    return out, time.perf_counter() - t0


def build_shards(fac: ShardFactory, specs):
    """Shards from (n, steps, seed, memory) specs; shards with the same (n, seed, memory) share one
    instance (its CCS, its per-step witnesses and memory on the device), built for the most steps
    any of them proves."""
    most = {}
    for n, steps, seed, memory in specs:
        key = (n, seed, bool(memory))
        most[key] = max(most.get(key, 0), steps)
    cache = {key: fac.make(key[0], st, key[1], memory=key[2]) for key, st in most.items()}
    return [cache[(n, seed, bool(memory))].with_steps(steps) for n, steps, seed, memory in specs]


def _proc_worker(pp_seed, specs, threads, tasks, results, share_rounds=True):
    """One prover process: builds the shards of `specs` on its own HIP runtime, warms `threads`
    contexts, reports ready, then proves shard indices from `tasks` on `threads` threads until a
    None per thread arrives, putting (index, proof) on `results`."""
    fac = ShardFactory(pp_seed)
    shards = build_shards(fac, specs)
    pool = ProverPool(threads)
    pool.warm(shards)
    if threads > 1 and share_rounds:
        import nightstream_amd as ns
        ns.set_shared_rounds(True)
    results.put(("ready", pool.created))

    def loop():
        while True:
            i = tasks.get()
            if i is None:
                return
            results.put((i, pool.prove(shards[i])))

    ths = [threading.Thread(target=loop) for _ in range(threads)]
    for th in ths:
        th.start()
    for th in ths:
        th.join()


class ProcessProver:
    """`procs` prover processes of `threads` contexts each over the same shard specs.  One process
    with many contexts tops out at about 100k GPU dispatches per second (launches and copies of
    short kernels); separate processes have their own HIP runtime and hardware queues and share
    only the device.  Give each few hardware queues (`hw_queues`, set in the children's
    environment): every mapped queue costs the GPU's queue scheduling time even when idle.
    Processes are started with `spawn` (a fresh interpreter: nothing of the parent's HIP state is
    inherited) and pull shard indices from one queue, so the load balances itself."""

    def __init__(self, pp_seed: bytes, specs, procs: int, threads: int, hw_queues: int = 0):
        import multiprocessing as mp
        import os
        ctx = mp.get_context("spawn")
        self.n = len(specs)
        self.procs, self.threads = procs, threads
        self.tasks, self.results = ctx.Queue(), ctx.Queue()
        self.ps = [ctx.Process(target=_proc_worker, args=(pp_seed, specs, threads, self.tasks, self.results, SHARE_ROUNDS),
                               daemon=True) for _ in range(procs)]
        old = os.environ.get("GPU_MAX_HW_QUEUES")
        if hw_queues:  # hardware queues per worker process (the children read it at HIP init)
            os.environ["GPU_MAX_HW_QUEUES"] = str(hw_queues)
        try:
            for p in self.ps:
                p.start()
        finally:
            if hw_queues:
                if old is None:
                    os.environ.pop("GPU_MAX_HW_QUEUES", None)
                else:
                    os.environ["GPU_MAX_HW_QUEUES"] = old
        self.created = 0
        for _ in range(procs):
            tag, created = self._get()
            if tag != "ready":
                raise RuntimeError("prover process failed to start")
            self.created += created

    def _get(self):
        import queue as q
        while True:
            try:
                return self.results.get(timeout=5.0)
            except q.Empty:
                dead = [p.exitcode for p in self.ps if p.exitcode not in (None, 0)]
                if dead:
                    raise RuntimeError(f"prover process exited with status {dead[0]}")

    def prove_all(self, order):
        """Prove every shard (in `order` of submission); returns (proofs by index, seconds)."""
        out = [None] * self.n
        t0 = time.perf_counter()
        for i in order:
            self.tasks.put(i)
        for _ in range(len(order)):
            i, proof = self._get()
            out[i] = proof
        return out, time.perf_counter() - t0

    def close(self):
        for _ in range(self.procs * self.threads):
            self.tasks.put(None)
        for p in self.ps:
            p.join(timeout=60)
        bad = [p.exitcode for p in self.ps if p.exitcode != 0]
        if bad:
            raise RuntimeError(f"prover process exited with status {bad[0]}")


@dataclasses.dataclass
class JobResult:
    job: int
    n: int
    steps: int
    arrival: float
    start: float
    finish: float
    proof: bytes

    @property
    def latency(self):
        return self.finish - self.arrival


def serve(jobs, rate: float, duration: float, workers: int, hbm_cap: float, seed: int = 0, pp_resident: float = 0.0,
          pool: ProverPool | None = None):
    """Open-loop Poisson arrivals at `rate` jobs/s for `duration` seconds, drawing jobs in order
    from `jobs` (a list of Shard, cycled).  A job starts when a worker is free and the summed
    footprint of running jobs plus its own stays within `hbm_cap` bytes (continuous admission:
    finished jobs free their slot immediately).  Returns the completed JobResult list; jobs that
    arrive after `duration` are not generated, and every admitted job runs to completion."""
    rng = np.random.default_rng(seed)
    arrivals = []
    t = 0.0
    while True:
        t += rng.exponential(1.0 / rate)
        if t > duration:
            break
        arrivals.append(t)
    pool = pool or ProverPool(workers)
    pending = queue.Queue()
    results = []
    res_lock = threading.Lock()
    cond = threading.Condition()
    state = {"mem": pp_resident, "free": workers}
    t0 = time.perf_counter()

    def worker():
        while True:
            item = pending.get()
            if item is None:
                return
            idx, sh, arr, fp = item
            start = time.perf_counter() - t0
            proof = pool.prove(sh)
            fin = time.perf_counter() - t0
            with res_lock:
                results.append(JobResult(idx, sh.n, sh.steps, arr, start, fin, proof))
            with cond:
                state["mem"] -= fp
                state["free"] += 1
                cond.notify_all()

    threads = [threading.Thread(target=worker, daemon=True) for _ in range(workers)]
    with shared_rounds(workers > 1):
        for th in threads:
            th.start()
        for i, arr in enumerate(arrivals):
            now = time.perf_counter() - t0
            if arr > now:
                time.sleep(arr - now)
            sh = jobs[i % len(jobs)]
            fp = footprint_bytes(sh.n) + (mem_footprint_bytes(sh.n) if sh.mem is not None else 0)
            with cond:
                # FIFO admission: wait for a slot and for HBM headroom
                while state["free"] == 0 or state["mem"] + fp > hbm_cap:
                    cond.wait()
                state["free"] -= 1
                state["mem"] += fp
            pending.put((i, sh, arr, fp))
        for _ in threads:
            pending.put(None)
        for th in threads:
            th.join()
    results.sort(key=lambda r: r.job)
    return results
