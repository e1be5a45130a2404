"""ctypes bindings to the CPU restatement (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's CPU
baseline.  The MI355X product library never loads this.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
Q = 0xFFFFFFFF00000001
D = 54

_lib = None


class KX(C.Structure):
    """K element passed by value (struct { uint64_t c0, c1; })."""
    _fields_ = [("c0", C.c_uint64), ("c1", C.c_uint64)]


def build(force: bool = False) -> str:
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-C", HERE, "-B" if force else "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = C.CDLL(LIB_PATH)
        u64p = np.ctypeslib.ndpointer(np.uint64, flags="C_CONTIGUOUS")
        i8p = np.ctypeslib.ndpointer(np.int8, flags="C_CONTIGUOUS")
        i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
        u32p = np.ctypeslib.ndpointer(np.uint32, flags="C_CONTIGUOUS")
        L.or_fq_mul_batch.argtypes = [u64p, u64p, u64p, C.c_size_t]
        L.or_k_mul_batch.argtypes = [u64p, u64p, u64p, C.c_size_t]
        L.or_k_inv_batch.argtypes = [u64p, u64p, C.c_size_t]
        L.or_rq_mul.argtypes = [u64p, u64p, u64p]
        L.or_rot_step.argtypes = [u64p, u64p]
        L.or_chacha_ietf_block.argtypes = [u32p, C.c_uint32, u32p, C.c_int, u32p]
        L.or_chacha_block.argtypes = [u32p, C.c_uint64, C.c_uint64, C.c_int, u32p]
        L.or_p2_permute8.argtypes = [u64p]
        L.or_p2_digest_perm.restype = C.c_void_p
        L.or_p2_permute.argtypes = [C.c_void_p, u64p]
        L.or_decomp_b2_row_major.argtypes = [u64p, C.c_size_t, i8p]
        L.or_ajtai_setup.argtypes = [C.c_char_p, C.c_size_t, C.c_size_t, u64p]
        L.or_ajtai_commit.argtypes = [u64p, C.c_size_t, C.c_size_t, i8p, u64p]
        L.or_ccs_from_r1cs.restype = C.c_void_p
        L.or_ccs_from_r1cs.argtypes = [C.c_size_t, C.c_size_t] + [u64p, u32p, u64p] * 3
        L.or_ccs_free.argtypes = [C.c_void_p]
        L.or_ccs_digest.argtypes = [C.c_void_p, u64p]
        L.or_chi_table.argtypes = [u64p, C.c_size_t, u64p]
        L.or_rowstate_build.restype = C.c_void_p
        L.or_rowstate_build.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, u64p, u64p, u64p,
                                        KX, C.c_size_t, C.c_size_t, C.c_void_p]
        L.or_rowstate_free.argtypes = [C.c_void_p]
        L.or_rowstate_round_coeffs.restype = C.c_size_t
        L.or_rowstate_round_coeffs.argtypes = [C.c_void_p, u64p]
        L.or_rowstate_evals.argtypes = [C.c_void_p, u64p, C.c_size_t, u64p]
        L.or_rowstate_fold.argtypes = [C.c_void_p, KX]
        L.or_rowstate_len.restype = C.c_size_t
        L.or_rowstate_len.argtypes = [C.c_void_p]
        L.or_rowstate_export.restype = C.c_size_t
        L.or_rowstate_export.argtypes = [C.c_void_p, C.c_int, C.c_size_t, u64p, C.c_size_t]
        L.or_precompute_y.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, u64p, C.c_size_t, u64p]
        L.or_vjs.argtypes = [C.c_void_p, u64p, C.c_size_t, u64p]
        L.or_interpolate.argtypes = [u64p, u64p, C.c_size_t, u64p]
        L.or_rot_from_coeffs.argtypes = [i8p, u64p]
        L.or_rlc_zmix.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t, i32p]
        L.or_split_b.restype = C.c_int
        L.or_split_b.argtypes = [i32p, C.c_size_t, C.c_size_t, C.c_void_p, C.POINTER(C.c_int)]
        L.or_prover_new.restype = C.c_void_p
        L.or_prover_new.argtypes = [C.c_void_p, C.c_char_p, C.c_int]
        L.or_prover_free.argtypes = [C.c_void_p]
        L.or_prove_shard.restype = C.c_int
        L.or_prove_shard.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                     C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_double)]
        L.or_free.argtypes = [C.c_void_p]
        L.or_prove_shard_mem.restype = C.c_int
        L.or_prove_shard_mem.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                         C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t),
                                         C.POINTER(C.c_double)]
        L.or_prove_shard_mem_ob.restype = C.c_int
        L.or_prove_shard_mem_ob.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t, C.c_size_t,
                                            C.c_void_p, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(C.c_size_t),
                                            C.POINTER(C.c_double)]
        L.or_last_error.restype = C.c_char_p
        L.or_last_step_seconds.restype = C.c_double
        L.or_last_step_seconds.argtypes = [C.c_size_t]
        L.or_num_threads.restype = C.c_int
        L.or_tr_permutation_count.restype = C.c_uint64
        L.or_set_serial_where_serial.argtypes = [C.c_int]
        L.or_params_auto_r1cs.restype = C.c_int
        L.or_extension_check.restype = C.c_int
        _lib = L
    return _lib


def _ptrs(arrays):
    arr = (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])
    return arr


def _k2(x):
    return KX(int(x[0]), int(x[1]))


class Ccs:
    def __init__(self, inst):
        L = lib()
        self._keep = [inst.a, inst.b, inst.c]
        args = []
        for (rp, ci, v) in (inst.a, inst.b, inst.c):
            args += [np.ascontiguousarray(rp, np.uint64), np.ascontiguousarray(ci, np.uint32),
                     np.ascontiguousarray(v, np.uint64)]
        self._args = args
        self.ptr = L.or_ccs_from_r1cs(inst.n, inst.m, *args)
        self.n, self.m, self.t = inst.n, inst.m, 4

    def digest(self):
        out = np.zeros(4, np.uint64)
        lib().or_ccs_digest(self.ptr, out)
        return out

    def __del__(self):
        try:
            lib().or_ccs_free(self.ptr)
        except Exception:
            pass


def num_threads() -> int:
    return lib().or_num_threads()


def set_serial_where_serial(on: bool) -> None:
    """Run the loops the reference runs on one thread (Pi_RLC Z_mix, precompute_for_r) on one
    thread; the rest keeps every OpenMP thread (where the reference uses rayon)."""
    lib().or_set_serial_where_serial(1 if on else 0)


def p2_permute(state) -> np.ndarray:
    """Poseidon2 of the oracle: width 8 (transcript) or 16 (CCS digest); returns a copy."""
    st = np.ascontiguousarray(np.asarray(state, dtype=np.uint64)).copy()
    if st.size == 8:
        lib().or_p2_permute8(st)
    else:
        lib().or_p2_permute(lib().or_p2_digest_perm(), st)
    return st


def decomp(z: np.ndarray) -> np.ndarray:
    z = np.ascontiguousarray(z, np.uint64)
    Z = np.zeros(D * len(z), np.int8)
    lib().or_decomp_b2_row_major(z, len(z), Z)
    return Z.reshape(D, len(z))


def ajtai_setup(seed: bytes, kappa: int, m: int) -> np.ndarray:
    pp = np.zeros(kappa * m * D, np.uint64)
    lib().or_ajtai_setup(seed, kappa, m, pp)
    return pp


def ajtai_commit(pp: np.ndarray, kappa: int, m: int, Z: np.ndarray) -> np.ndarray:
    out = np.zeros(kappa * D, np.uint64)
    lib().or_ajtai_commit(np.ascontiguousarray(pp, np.uint64), kappa, m, np.ascontiguousarray(Z, np.int8).reshape(-1), out)
    return out


class RowState:
    """Pi_CCS row-phase oracle state (optimized_engine/oracle.rs RowStreamState)."""

    def __init__(self, ccs: Ccs, Zs, alpha, beta_a, beta_r, gamma, ell_d, ell_n, r_inputs=None):
        L = lib()
        self.Zs = [np.ascontiguousarray(z, np.int8).reshape(-1) for z in Zs]
        self._zp = _ptrs(self.Zs)
        a = np.ascontiguousarray(np.asarray(alpha, np.uint64).reshape(-1))
        ba = np.ascontiguousarray(np.asarray(beta_a, np.uint64).reshape(-1))
        br = np.ascontiguousarray(np.asarray(beta_r, np.uint64).reshape(-1))
        self._keep = [a, ba, br]
        rin = None
        if r_inputs is not None:
            rin = np.ascontiguousarray(np.asarray(r_inputs, np.uint64).reshape(-1))
            self._keep.append(rin)
        self.ptr = L.or_rowstate_build(ccs.ptr, self._zp, len(self.Zs), a, ba, br, _k2(gamma), ell_d, ell_n,
                                       None if rin is None else rin.ctypes.data)

    def coeffs(self):
        out = np.zeros(32, np.uint64)
        n = lib().or_rowstate_round_coeffs(self.ptr, out)
        return out[:2 * n].reshape(n, 2)

    def evals(self, xs):
        xs = np.ascontiguousarray(np.asarray(xs, np.uint64).reshape(-1))
        out = np.zeros_like(xs)
        lib().or_rowstate_evals(self.ptr, xs, len(xs) // 2, out)
        return out.reshape(-1, 2)

    def fold(self, r):
        lib().or_rowstate_fold(self.ptr, _k2(r))

    def length(self):
        return lib().or_rowstate_len(self.ptr)

    def export(self, which, idx=0):
        cap = self.length() * D
        out = np.zeros(2 * cap, np.uint64)
        n = lib().or_rowstate_export(self.ptr, which, idx, out, cap)
        return out[:2 * n].reshape(n, 2)

    def __del__(self):
        try:
            lib().or_rowstate_free(self.ptr)
        except Exception:
            pass


def precompute_y(ccs: Ccs, Zs, r, ell_n):
    Zs = [np.ascontiguousarray(z, np.int8).reshape(-1) for z in Zs]
    r = np.ascontiguousarray(np.asarray(r, np.uint64).reshape(-1))
    out = np.zeros(len(Zs) * ccs.t * D * 2, np.uint64)
    lib().or_precompute_y(ccs.ptr, _ptrs(Zs), len(Zs), r, ell_n, out)
    return out.reshape(len(Zs), ccs.t, D, 2)


def vjs(ccs: Ccs, r, ell_n):
    r = np.ascontiguousarray(np.asarray(r, np.uint64).reshape(-1))
    out = np.zeros(ccs.t * ccs.m * 2, np.uint64)
    lib().or_vjs(ccs.ptr, r, ell_n, out)
    return out.reshape(ccs.t, ccs.m, 2)


def rot_from_coeffs(a) -> np.ndarray:
    a = np.ascontiguousarray(np.asarray(a, np.int8))
    out = np.zeros(D * D, np.uint64)
    lib().or_rot_from_coeffs(a, out)
    return out.reshape(D, D)


def rlc_zmix(rhos, Zs, m):
    rhos = [np.ascontiguousarray(r, np.uint64).reshape(-1) for r in rhos]
    Zs = [np.ascontiguousarray(z, np.int8).reshape(-1) for z in Zs]
    out = np.zeros(D * m, np.int32)
    lib().or_rlc_zmix(_ptrs(rhos), _ptrs(Zs), len(Zs), m, out)
    return out.reshape(D, m)


def split_b(zmix: np.ndarray, k: int):
    zmix = np.ascontiguousarray(zmix, np.int32).reshape(-1)
    outs = [np.zeros(len(zmix), np.int8) for _ in range(k)]
    nz = (C.c_int * k)()
    rc = lib().or_split_b(zmix, len(zmix), k, _ptrs(outs), nz)
    if rc != 0:
        raise ValueError("DEC split out of range")
    return outs, [bool(x) for x in nz]


class LutInstC(C.Structure):
    _fields_ = [("k", C.c_size_t), ("d", C.c_size_t), ("n_side", C.c_size_t), ("steps", C.c_size_t),
                ("ell", C.c_size_t), ("table", C.c_void_p)]


class MemInstC(C.Structure):
    _fields_ = [("k", C.c_size_t), ("d", C.c_size_t), ("n_side", C.c_size_t), ("steps", C.c_size_t),
                ("ell", C.c_size_t), ("sparse", C.c_int), ("ninit", C.c_size_t), ("init_addrs", C.c_void_p),
                ("init_vals", C.c_void_p)]


class StepMemC(C.Structure):
    _fields_ = [("n_lut", C.c_size_t), ("luts", C.c_void_p), ("n_mem", C.c_size_t), ("mems", C.c_void_p)]


def step_mem_structs(mems, keep: list):
    """ctypes mirrors of per-step memory instances (objects with .luts / .mems as in
    nightstream_amd.bus_workload.StepMem); arrays they point into are appended to `keep`."""
    arr = (StepMemC * len(mems))()
    for s, sm in enumerate(mems):
        luts = (LutInstC * max(1, len(sm.luts)))()
        for i, lt in enumerate(sm.luts):
            tab = np.ascontiguousarray(lt.table, np.uint64)
            keep.append(tab)
            luts[i] = LutInstC(lt.k, lt.d, lt.n_side, lt.steps, lt.ell, tab.ctypes.data)
        ms = (MemInstC * max(1, len(sm.mems)))()
        for i, mi in enumerate(sm.mems):
            if mi.init is None:
                ms[i] = MemInstC(mi.k, mi.d, mi.n_side, mi.steps, mi.ell, 0, 0, None, None)
            else:
                a = np.array([x for x, _ in mi.init] or [0], np.uint64)
                v = np.array([y for _, y in mi.init] or [0], np.uint64)
                keep += [a, v]
                ms[i] = MemInstC(mi.k, mi.d, mi.n_side, mi.steps, mi.ell, 1, len(mi.init), a.ctypes.data, v.ctypes.data)
        keep += [luts, ms]
        arr[s] = StepMemC(len(sm.luts), C.cast(luts, C.c_void_p), len(sm.mems), C.cast(ms, C.c_void_p))
    keep.append(arr)
    return arr


class ObConfigC(C.Structure):
    """or_ob_config: the output-binding claims (ProgramIO) and the final memory state."""
    _fields_ = [("mem_idx", C.c_size_t), ("num_bits", C.c_size_t), ("n_claims", C.c_size_t),
                ("addrs", C.c_void_p), ("values", C.c_void_p), ("final_state", C.c_void_p)]


def ob_struct(ob, keep):
    """ob: dict(mem_idx, num_bits, claims=[(addr, value)], final_state=array of 2^num_bits)."""
    claims = sorted(ob["claims"])
    a = np.array([x for x, _ in claims] or [0], np.uint64)
    v = np.array([y for _, y in claims] or [0], np.uint64)
    fs = np.ascontiguousarray(ob["final_state"], np.uint64)
    if fs.size != 1 << ob["num_bits"]:
        raise ValueError("final_state must hold 2^num_bits values")
    keep += [a, v, fs]
    c = ObConfigC(ob["mem_idx"], ob["num_bits"], len(claims), a.ctypes.data, v.ctypes.data, fs.ctypes.data)
    keep.append(c)
    return c


class Prover:
    """CPU shard prover (fold loop restatement); proves with the canonical proof encoding."""

    def __init__(self, inst, pp_seed: bytes, n_threads: int = 0):
        self.ccs = Ccs(inst)
        self.ptr = lib().or_prover_new(self.ccs.ptr, pp_seed, n_threads)
        if not self.ptr:
            raise RuntimeError(lib().or_last_error().decode())

    def prove(self, zs, m_in=1, mem=None, ob=None):
        """mem: per-step memory instances (Route A over the shared CPU bus), or None.
        ob: output binding on the last step (see ob_struct), or None."""
        zs = [np.ascontiguousarray(z, np.uint64) for z in zs]
        out = C.POINTER(C.c_uint8)()
        n = C.c_size_t()
        sec = C.c_double()
        keep = []
        marr = step_mem_structs(mem, keep) if mem is not None else None
        obs = C.byref(ob_struct(ob, keep)) if ob is not None else None
        rc = lib().or_prove_shard_mem_ob(self.ptr, _ptrs(zs), None if marr is None else C.cast(marr, C.c_void_p),
                                         len(zs), m_in, obs, C.byref(out), C.byref(n), C.byref(sec))
        if rc != 0:
            raise RuntimeError(f"oracle prove failed ({rc}): {lib().or_last_error().decode()}")
        data = bytes(C.cast(out, C.POINTER(C.c_uint8 * n.value)).contents)
        lib().or_free(out)
        self.step_seconds = [lib().or_last_step_seconds(i) for i in range(len(zs))]
        return data, sec.value

    def __del__(self):
        try:
            lib().or_prover_free(self.ptr)
        except Exception:
            pass


# ---------------- Twist / Shout oracles (oracle/twist.c) ----------------
class SparseIn(C.Structure):
    _fields_ = [("nnz", C.c_size_t), ("idx", C.POINTER(C.c_uint64)), ("vals", C.POINTER(C.c_uint64))]


def _kbuf(ks):
    """K values as a flat uint64 buffer [c0, c1, ...]: an (n, 2) uint64 array is used as it is,
    any other sequence of (c0, c1) pairs is converted element by element."""
    if isinstance(ks, np.ndarray) and ks.dtype == np.uint64 and ks.ndim == 2 and ks.shape[1] == 2:
        return np.ascontiguousarray(ks).reshape(-1) if len(ks) else np.zeros(2, np.uint64)
    return np.array([v for k in ks for v in (int(k[0]), int(k[1]))] or [0, 0], dtype=np.uint64)


class _Keep:
    def __init__(self):
        self.items = []

    def sparse(self, col):
        """col: (idx array, list of K) -> SparseIn (keeps the buffers alive)."""
        idx = np.ascontiguousarray(np.asarray(col[0], np.uint64).reshape(-1)) if len(col[0]) else np.zeros(1, np.uint64)
        vals = _kbuf(col[1])
        self.items += [idx, vals]
        return SparseIn(len(col[0]), idx.ctypes.data_as(C.POINTER(C.c_uint64)),
                        vals.ctypes.data_as(C.POINTER(C.c_uint64)))

    def sparse_arr(self, cols):
        arr = (SparseIn * max(1, len(cols)))(*[self.sparse(c) for c in cols])
        self.items.append(arr)
        return arr

    def ks(self, ks):
        a = _kbuf(ks)
        self.items.append(a)
        return a.ctypes.data_as(C.c_void_p)


def _ts_lib():
    L = lib()
    if getattr(L, "_ts_bound", False):
        return L
    vp, sz = C.c_void_p, C.c_size_t
    KP = C.POINTER(KX)
    L.or_ts_product.restype = vp
    L.or_ts_product.argtypes = [vp, sz, sz, sz]
    L.or_ts_shout_value.restype = vp
    L.or_ts_shout_value.argtypes = [vp, sz, SparseIn, SparseIn, KP]
    L.or_ts_index_adapter.restype = vp
    L.or_ts_index_adapter.argtypes = [vp, sz, SparseIn, vp, sz, vp, KP]
    L.or_ts_bitness.restype = vp
    L.or_ts_bitness.argtypes = [vp, sz, vp, vp, sz]
    L.or_ts_twist_read.restype = vp
    L.or_ts_twist_read.argtypes = [vp, sz, SparseIn, SparseIn, vp, SparseIn, SparseIn, vp, sz, vp, KX]
    L.or_ts_twist_write.restype = vp
    L.or_ts_twist_write.argtypes = [vp, sz, SparseIn, SparseIn, SparseIn, vp, sz, vp, KX]
    L.or_ts_val_eval.restype = vp
    L.or_ts_val_eval.argtypes = [sz, vp, sz, SparseIn, SparseIn, vp, vp, KP]
    L.or_ts_total_inc.restype = vp
    L.or_ts_total_inc.argtypes = [sz, vp, sz, SparseIn, SparseIn, vp, KP]
    L.or_ts_read_addr.restype = vp
    L.or_ts_read_addr.argtypes = [vp, vp, sz, vp, sz, SparseIn, SparseIn, vp, SparseIn, vp, SparseIn, sz]
    L.or_ts_write_addr.restype = vp
    L.or_ts_write_addr.argtypes = [vp, vp, sz, vp, sz, SparseIn, SparseIn, vp, SparseIn, sz]
    L.or_ts_addr_lookup.restype = vp
    L.or_ts_addr_lookup.argtypes = [vp, sz, SparseIn, vp, sz, vp, sz, KP]
    L.or_ts_evals.argtypes = [vp, vp, sz, vp]
    L.or_ts_fold.argtypes = [vp, KX]
    L.or_ts_num_rounds.restype = sz
    L.or_ts_num_rounds.argtypes = [vp]
    L.or_ts_degree_bound.restype = sz
    L.or_ts_degree_bound.argtypes = [vp]
    L.or_ts_free.argtypes = [vp]
    L._ts_bound = True
    return L


class TsOracle:
    """A Twist/Shout round oracle of the CPU restatement.  Columns are (idx, [K]) sparse pairs."""

    def __init__(self, kind, *args):
        L = _ts_lib()
        k = _Keep()
        self._keep = k
        claim = KX(0, 0)
        cp = C.byref(claim)
        if kind == "product":
            factors, degree_bound = args
            bufs = [_kbuf(f) for f in factors]
            k.items += bufs
            ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
            k.items.append(ptrs)
            self.h = L.or_ts_product(ptrs, len(factors), len(factors[0]), degree_bound)
        elif kind == "shout_value":
            r_cycle, has_lookup, val = args
            self.h = L.or_ts_shout_value(k.ks(r_cycle), len(r_cycle), k.sparse(has_lookup), k.sparse(val), cp)
        elif kind == "index_adapter":
            r_cycle, has_lookup, addr_bits, r_addr = args
            self.h = L.or_ts_index_adapter(k.ks(r_cycle), len(r_cycle), k.sparse(has_lookup), k.sparse_arr(addr_bits),
                                           len(addr_bits), k.ks(r_addr), cp)
        elif kind == "bitness":
            r_cycle, cols, weights = args
            self.h = L.or_ts_bitness(k.ks(r_cycle), len(r_cycle), k.sparse_arr(cols), k.ks(weights), len(cols))
        elif kind == "twist_read":
            r_cycle, has_read, rv, ra_bits, has_write, inc, wa_bits, r_addr, init_at = args
            self.h = L.or_ts_twist_read(k.ks(r_cycle), len(r_cycle), k.sparse(has_read), k.sparse(rv),
                                        k.sparse_arr(ra_bits), k.sparse(has_write), k.sparse(inc),
                                        k.sparse_arr(wa_bits), len(r_addr), k.ks(r_addr), KX(*init_at))
        elif kind == "twist_write":
            r_cycle, has_write, wv, inc, wa_bits, r_addr, init_at = args
            self.h = L.or_ts_twist_write(k.ks(r_cycle), len(r_cycle), k.sparse(has_write), k.sparse(wv),
                                         k.sparse(inc), k.sparse_arr(wa_bits), len(r_addr), k.ks(r_addr), KX(*init_at))
        elif kind == "val_eval":
            ell_n, wa_bits, has_write, inc, r_addr, r_time = args
            self.h = L.or_ts_val_eval(ell_n, k.sparse_arr(wa_bits), len(wa_bits), k.sparse(has_write), k.sparse(inc),
                                      k.ks(r_addr), k.ks(r_time), cp)
        elif kind == "total_inc":
            ell_n, wa_bits, has_write, inc, r_addr = args
            self.h = L.or_ts_total_inc(ell_n, k.sparse_arr(wa_bits), len(wa_bits), k.sparse(has_write),
                                       k.sparse(inc), k.ks(r_addr), cp)
        elif kind == "read_addr":
            init, r_cycle, has_read, rv, ra_bits, has_write, wa_bits, inc = args
            ia = np.array([a for a, _ in init] or [0], np.uint64)
            k.items.append(ia)
            self.h = L.or_ts_read_addr(ia.ctypes.data, k.ks([v for _, v in init]), len(init), k.ks(r_cycle),
                                       len(r_cycle), k.sparse(has_read), k.sparse(rv), k.sparse_arr(ra_bits),
                                       k.sparse(has_write), k.sparse_arr(wa_bits), k.sparse(inc), len(ra_bits))
        elif kind == "write_addr":
            init, r_cycle, has_write, wv, wa_bits, inc = args
            ia = np.array([a for a, _ in init] or [0], np.uint64)
            k.items.append(ia)
            self.h = L.or_ts_write_addr(ia.ctypes.data, k.ks([v for _, v in init]), len(init), k.ks(r_cycle),
                                        len(r_cycle), k.sparse(has_write), k.sparse(wv), k.sparse_arr(wa_bits),
                                        k.sparse(inc), len(wa_bits))
        elif kind == "addr_lookup":
            addr_bits, has_lookup, table, r_cycle = args
            self.h = L.or_ts_addr_lookup(k.sparse_arr(addr_bits), len(addr_bits), k.sparse(has_lookup), k.ks(table),
                                         len(table), k.ks(r_cycle), len(r_cycle), cp)
        else:
            raise ValueError(kind)
        if not self.h:
            raise RuntimeError("oracle construction failed")
        self.claim = (claim.c0, claim.c1)

    def num_rounds(self):
        return _ts_lib().or_ts_num_rounds(self.h)

    def degree_bound(self):
        return _ts_lib().or_ts_degree_bound(self.h)

    def evals(self, xs):
        xb = _kbuf(xs)
        out = np.zeros(2 * len(xs), np.uint64)
        _ts_lib().or_ts_evals(self.h, xb.ctypes.data, len(xs), out.ctypes.data)
        o = [int(v) for v in out]
        return [(o[2 * i], o[2 * i + 1]) for i in range(len(xs))]

    def fold(self, r):
        _ts_lib().or_ts_fold(self.h, KX(int(r[0]), int(r[1])))

    def __del__(self):
        try:
            _ts_lib().or_ts_free(self.h)
        except Exception:
            pass


# ---------------- shared CPU bus (oracle/bus.c) ----------------
def bus_decode(Z):
    Z = np.ascontiguousarray(Z, np.int8)
    m = Z.shape[1]
    out = np.zeros(m, np.uint64)
    L = lib()
    L.or_bus_decode.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    L.or_bus_decode(Z.ctypes.data, m, out.ctypes.data)
    return out


def bus_time_col(z, m_in, bus_base, chunk, bus_cols, col_id, steps_len, pow2_cycle):
    z = np.ascontiguousarray(z, np.uint64)
    idx = np.zeros(chunk + 1, np.uint64)
    vals = np.zeros(2 * chunk + 2, np.uint64)
    L = lib()
    L.or_bus_time_col.restype = C.c_long
    L.or_bus_time_col.argtypes = [C.c_void_p] + [C.c_size_t] * 8 + [C.c_void_p, C.c_void_p]
    n = L.or_bus_time_col(z.ctypes.data, len(z), m_in, bus_base, chunk, bus_cols, col_id, steps_len, pow2_cycle,
                          idx.ctypes.data, vals.ctypes.data)
    if n < 0:
        raise ValueError("bus time column: invalid layout")
    return idx[:n].copy(), [(int(vals[2 * i]), int(vals[2 * i + 1])) for i in range(n)]


def bus_openings(Z, m_in, bus_base, chunk, bus_cols, r):
    Z = np.ascontiguousarray(Z, np.int8)
    m = Z.shape[1]
    rb = np.array([v for k in r for v in (int(k[0]), int(k[1]))], np.uint64)
    y = np.zeros(bus_cols * 64 * 2, np.uint64)
    ys = np.zeros(bus_cols * 2, np.uint64)
    L = lib()
    L.or_bus_openings.restype = C.c_int
    L.or_bus_openings.argtypes = [C.c_void_p] + [C.c_size_t] * 5 + [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    rc = L.or_bus_openings(Z.ctypes.data, m, m_in, bus_base, chunk, bus_cols, rb.ctypes.data, len(r), y.ctypes.data,
                           ys.ctypes.data)
    if rc != 0:
        raise ValueError("bus openings: invalid layout")
    yl = [int(v) for v in y]
    ysl = [int(v) for v in ys]
    return ([[(yl[2 * (c * 64 + i)], yl[2 * (c * 64 + i) + 1]) for i in range(64)] for c in range(bus_cols)],
            [(ysl[2 * c], ysl[2 * c + 1]) for c in range(bus_cols)])
