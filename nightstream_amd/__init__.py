"""MI355X-native backend for the Nightstream Neo folding hot path.

The compute lives in libnightstream.so (hand-written HIP for gfx950 plus a C++ host fold loop)
behind the C ABI in include/nightstream.h; this module binds it with ctypes and mirrors the
reference's entry points (neo-fold FoldingSession, neo-ajtai commit, the Pi_CCS RoundOracle).
There is no CPU fallback: every call raises NsError when the library or the GPU is missing.
"""
from __future__ import annotations

import ctypes as C
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libnightstream.so")
Q = 0xFFFFFFFF00000001
D = 54
KAPPA = 16

EXPORTS = [
    "ns_last_error", "ns_device_count", "ns_free", "ns_device_alloc", "ns_device_free", "ns_memcpy_h2d",
    "ns_memcpy_d2h", "ns_device_synchronize", "ns_params_goldilocks_auto_r1cs_ccs", "ns_ccs_from_r1cs",
    "ns_ccs_digest", "ns_ccs_free", "ns_pp_setup_seeded", "ns_pp_from_coeffs", "ns_pp_coeffs", "ns_pp_free",
    "ns_decomp_b", "ns_ajtai_commit_batch", "ns_ajtai_commit_batch_device", "ns_ajtai_force_variant",
    "ns_set_shared_rounds", "ns_shared_round_stats", "ns_ccs_oracle_new", "ns_oracle_evals", "ns_oracle_fold",
    "ns_oracle_num_rounds", "ns_oracle_degree_bound", "ns_oracle_y_eval", "ns_oracle_free", "ns_rlc_zmix",
    "ns_dec_split", "ns_dec_split_device_noalloc", "ns_prover_new", "ns_prove_shard_mem", "ns_shard_verify", "ns_prove_shard_ob", "ns_shard_verify_ob", "ns_rlc", "ns_dec", "ns_prover_rebind", "ns_prove_shard", "ns_prove_shard_device", "ns_prover_set_timing",
    "ns_prover_kernel_stats", "ns_prover_phase_stats", "ns_prover_phase_perms", "ns_prover_free",
    "ns_poseidon2_permute",
    "ns_ts_product_new", "ns_ts_shout_value_new", "ns_ts_index_adapter_new", "ns_ts_bitness_new",
    "ns_ts_twist_read_new", "ns_ts_twist_write_new", "ns_ts_val_eval_new", "ns_ts_total_inc_new",
    "ns_ts_read_addr_new", "ns_ts_write_addr_new", "ns_ts_addr_lookup_new", "ns_ts_evals", "ns_ts_fold",
    "ns_ts_batch_round", "ns_ts_batch_fold",
    "ns_ts_num_rounds", "ns_ts_degree_bound", "ns_ts_free", "ns_bus_decode", "ns_bus_time_cols", "ns_bus_openings",
]


class NsError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


_lib = None
_lib_lock = threading.Lock()


def lib():
    """Load libnightstream.so (raises if it has not been built).  Thread-safe: the first caller
    loads and annotates the one library object every later call returns."""
    global _lib
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is None:
            _lib = _load()
    return _lib


def _load():
    if not os.path.exists(LIB_PATH):
        raise NsError(-10, f"{LIB_PATH} is missing: run `python -m nightstream_amd.build`")
    L = C.CDLL(LIB_PATH)
    vp, sz, u64p = C.c_void_p, C.c_size_t, C.POINTER(C.c_uint64)
    L.ns_last_error.restype = C.c_char_p
    L.ns_device_count.restype = C.c_int
    L.ns_free.argtypes = [vp]
    for name in EXPORTS:
        fn = getattr(L, name)
        if name not in ("ns_last_error", "ns_device_count", "ns_free", "ns_oracle_num_rounds",
                        "ns_oracle_degree_bound", "ns_ts_num_rounds", "ns_ts_degree_bound") \
                and not name.endswith("_free"):
            fn.restype = C.c_int
    L.ns_oracle_num_rounds.restype = sz
    L.ns_oracle_degree_bound.restype = sz
    L.ns_oracle_num_rounds.argtypes = [vp]
    L.ns_oracle_degree_bound.argtypes = [vp]
    for n in ("ns_ccs_free", "ns_pp_free", "ns_oracle_free", "ns_prover_free"):
        getattr(L, n).argtypes = [vp]
        getattr(L, n).restype = None
    L.ns_ccs_from_r1cs.argtypes = [sz, sz] + [vp] * 9 + [C.POINTER(vp)]
    L.ns_ccs_digest.argtypes = [vp, vp]
    L.ns_pp_setup_seeded.argtypes = [C.c_char_p, sz, C.c_int, C.POINTER(vp)]
    L.ns_pp_from_coeffs.argtypes = [vp, sz, C.c_int, C.POINTER(vp)]
    L.ns_pp_coeffs.argtypes = [vp, vp]
    L.ns_decomp_b.argtypes = [vp, sz, C.c_uint32, vp]
    L.ns_ajtai_commit_batch.argtypes = [vp, vp, sz, vp]
    L.ns_ajtai_commit_batch_device.argtypes = [vp, vp, sz, vp]
    L.ns_ajtai_force_variant.argtypes = [C.c_int]
    L.ns_set_shared_rounds.argtypes = [C.c_int]
    L.ns_shared_round_stats.argtypes = [C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
    L.ns_ccs_oracle_new.argtypes = [vp, vp, sz, vp, vp, vp, vp, sz, sz, vp, C.POINTER(vp)]
    L.ns_oracle_evals.argtypes = [vp, vp, sz, vp]
    L.ns_oracle_fold.argtypes = [vp, vp]
    L.ns_oracle_y_eval.argtypes = [vp, vp, vp]
    L.ns_rlc_zmix.argtypes = [vp, vp, sz, sz, vp]
    L.ns_dec_split.argtypes = [vp, sz, sz, vp, vp]
    L.ns_dec_split_device_noalloc.argtypes = [vp, sz, sz, vp, vp]
    L.ns_prover_new.argtypes = [vp, vp, C.POINTER(vp)]
    L.ns_prover_rebind.argtypes = [vp, vp, vp]
    L.ns_device_alloc.argtypes = [sz, C.POINTER(vp)]
    L.ns_device_free.argtypes = [vp]
    L.ns_memcpy_h2d.argtypes = [vp, vp, sz]
    L.ns_memcpy_d2h.argtypes = [vp, vp, sz]
    L.ns_prove_shard.argtypes = [vp, vp, sz, sz, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(sz),
                                 C.POINTER(C.c_double)]
    L.ns_prove_shard_device.argtypes = L.ns_prove_shard.argtypes
    L.ns_prove_shard_mem.argtypes = [vp, vp, C.c_int, sz, sz, vp, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(sz),
                                     C.POINTER(C.c_double)]
    L.ns_shard_verify.argtypes = [vp, vp, sz, sz, vp, vp, vp, sz]
    L.ns_prove_shard_ob.argtypes = [vp, vp, C.c_int, sz, sz, vp, vp, C.POINTER(C.POINTER(C.c_uint8)), C.POINTER(sz),
                                    C.POINTER(C.c_double)]
    L.ns_shard_verify_ob.argtypes = [vp, vp, sz, sz, vp, vp, vp, vp, sz]
    L.ns_rlc.argtypes = [vp, vp, sz, vp, vp, vp, vp]
    L.ns_dec.argtypes = [vp, vp, vp, sz, vp, sz, vp, vp]
    L.ns_prover_set_timing.argtypes = [vp, C.c_int]
    L.ns_prover_kernel_stats.argtypes = [vp, vp, vp, sz]
    L.ns_prover_phase_stats.argtypes = [vp, vp, sz]
    L.ns_prover_phase_perms.argtypes = [vp, vp, sz]
    L.ns_poseidon2_permute.argtypes = [vp, sz, C.c_int]
    L.ns_params_goldilocks_auto_r1cs_ccs.argtypes = [sz, vp]
    return L


def _check(rc: int):
    if rc != 0:
        raise NsError(rc, lib().ns_last_error().decode(errors="replace"))


class DeviceBuffer:
    """HBM allocation on the library's own HIP runtime (freed with the object)."""

    def __init__(self, nbytes: int):
        p = C.c_void_p()
        _check(lib().ns_device_alloc(nbytes, C.byref(p)))
        self.ptr = p.value
        self.nbytes = nbytes

    @classmethod
    def from_array(cls, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _check(lib().ns_memcpy_h2d(b.ptr, a.ctypes.data, a.nbytes))
        return b

    def to_array(self, dtype, count) -> np.ndarray:
        out = np.empty(count, dtype)
        _check(lib().ns_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes))
        return out

    def data_ptr(self) -> int:
        return self.ptr

    def __del__(self):
        try:
            if getattr(self, "ptr", None):
                lib().ns_device_free(self.ptr)
                self.ptr = None
        except Exception:
            pass



def _ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _ptr_array(arrays):
    return (C.c_void_p * len(arrays))(*[a.ctypes.data for a in arrays])


def device_count() -> int:
    return lib().ns_device_count()


def device_synchronize():
    _check(lib().ns_device_synchronize())


def poseidon2_permute(state, force_scalar: bool = False) -> np.ndarray:
    """The host Poseidon2 permutation (width 8: transcript, width 16: CCS digest) of a state of
    canonical Goldilocks elements; returns the permuted copy."""
    st = np.ascontiguousarray(np.asarray(state, dtype=np.uint64)).copy()
    _check(lib().ns_poseidon2_permute(_ptr(st), st.size, 1 if force_scalar else 0))
    return st


def params_goldilocks_auto_r1cs_ccs(n_rows: int) -> dict:
    out = np.zeros(11, np.uint64)
    _check(lib().ns_params_goldilocks_auto_r1cs_ccs(n_rows, _ptr(out)))
    keys = ["q", "eta", "d", "kappa", "m", "b", "k_rho", "B", "T", "s", "lambda"]
    return {k: int(v) for k, v in zip(keys, out)}


class Ccs:
    """Identity-first CCS from R1CS CSR matrices (neo-ccs r1cs_to_ccs)."""

    def __init__(self, inst):
        self.n, self.m = inst.n, inst.m
        arrs = []
        for (rp, ci, v) in (inst.a, inst.b, inst.c):
            arrs += [np.ascontiguousarray(rp, np.uint64), np.ascontiguousarray(ci, np.uint32),
                     np.ascontiguousarray(v, np.uint64)]
        self._keep = arrs
        h = C.c_void_p()
        _check(lib().ns_ccs_from_r1cs(inst.n, inst.m, *[_ptr(a) for a in arrs], C.byref(h)))
        self.h = h

    def digest(self) -> np.ndarray:
        out = np.zeros(4, np.uint64)
        _check(lib().ns_ccs_digest(self.h, _ptr(out)))
        return out

    def __del__(self):
        if getattr(self, "h", None):
            lib().ns_ccs_free(self.h)
            self.h = None


class PublicParams:
    """Ajtai PP resident in HBM (neo_ajtai::setup from a ChaCha8 seed)."""

    def __init__(self, m: int, seed: bytes | None = None, coeffs: np.ndarray | None = None, keep_coeffs=False):
        self.m = m
        h = C.c_void_p()
        if coeffs is not None:
            c = np.ascontiguousarray(coeffs, np.uint64)
            _check(lib().ns_pp_from_coeffs(_ptr(c), m, int(keep_coeffs), C.byref(h)))
        else:
            assert seed is not None and len(seed) == 32
            _check(lib().ns_pp_setup_seeded(seed, m, int(keep_coeffs), C.byref(h)))
        self.h = h

    def coeffs(self) -> np.ndarray:
        out = np.zeros(KAPPA * self.m * D, np.uint64)
        _check(lib().ns_pp_coeffs(self.h, _ptr(out)))
        return out

    def commit(self, Zs) -> np.ndarray:
        """SModuleHomomorphism::commit for a batch of D x m digit matrices -> [n, kappa*D]."""
        Zs = [np.ascontiguousarray(z, np.int8).reshape(-1) for z in Zs]
        out = np.zeros(len(Zs) * KAPPA * D, np.uint64)
        _check(lib().ns_ajtai_commit_batch(self.h, _ptr_array(Zs), len(Zs), _ptr(out)))
        return out.reshape(len(Zs), KAPPA * D)

    def commit_device(self, bufs) -> np.ndarray:
        """commit for D x m digit matrices already in HBM (DeviceBuffer objects) -> [n, kappa*D]."""
        ptrs = (C.c_void_p * max(1, len(bufs)))(*[b.ptr for b in bufs])
        out = np.zeros(max(1, len(bufs)) * KAPPA * D, np.uint64)
        _check(lib().ns_ajtai_commit_batch_device(self.h, ptrs, len(bufs), _ptr(out)))
        return out[:len(bufs) * KAPPA * D].reshape(len(bufs), KAPPA * D)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ns_pp_free(self.h)
            self.h = None


def decomp_b(z: np.ndarray, b: int = 2) -> np.ndarray:
    z = np.ascontiguousarray(z, np.uint64)
    Z = np.zeros(D * len(z), np.int8)
    _check(lib().ns_decomp_b(_ptr(z), len(z), b, _ptr(Z)))
    return Z.reshape(D, len(z))


def _kflat(xs) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(xs, np.uint64).reshape(-1))


class CcsOracle:
    """Pi_CCS RoundOracle on the GPU (evals_at / fold / num_rounds / degree_bound)."""

    def __init__(self, ccs: Ccs, Zs, alpha, beta_a, beta_r, gamma, ell_d, ell_n, r_inputs=None):
        self.Zs = [np.ascontiguousarray(z, np.int8).reshape(-1) for z in Zs]
        self._k = [_kflat(alpha), _kflat(beta_a), _kflat(beta_r), _kflat(gamma)]
        rin = None if r_inputs is None else _kflat(r_inputs)
        self._rin = rin
        h = C.c_void_p()
        _check(lib().ns_ccs_oracle_new(ccs.h, _ptr_array(self.Zs), len(self.Zs), *[_ptr(a) for a in self._k],
                                       ell_d, ell_n, None if rin is None else _ptr(rin), C.byref(h)))
        self.h = h
        self.ccs = ccs

    def num_rounds(self) -> int:
        return lib().ns_oracle_num_rounds(self.h)

    def degree_bound(self) -> int:
        return lib().ns_oracle_degree_bound(self.h)

    def evals_at(self, xs) -> np.ndarray:
        xs = _kflat(xs)
        out = np.zeros_like(xs)
        _check(lib().ns_oracle_evals(self.h, _ptr(xs), len(xs) // 2, _ptr(out)))
        return out.reshape(-1, 2)

    def fold(self, r):
        r = _kflat(r)
        _check(lib().ns_oracle_fold(self.h, _ptr(r)))

    def y_eval(self, r) -> np.ndarray:
        r = _kflat(r)
        out = np.zeros(len(self.Zs) * 4 * D * 2, np.uint64)
        _check(lib().ns_oracle_y_eval(self.h, _ptr(r), _ptr(out)))
        return out.reshape(len(self.Zs), 4, D, 2)

    def __del__(self):
        if getattr(self, "h", None):
            lib().ns_oracle_free(self.h)
            self.h = None


def rlc_zmix(rho_coeffs, Zs, m: int) -> np.ndarray:
    rc = np.ascontiguousarray(np.asarray(rho_coeffs, np.int8).reshape(-1))
    Zs = [np.ascontiguousarray(z, np.int8).reshape(-1) for z in Zs]
    out = np.zeros(D * m, np.int32)
    _check(lib().ns_rlc_zmix(_ptr(rc), _ptr_array(Zs), len(Zs), m, _ptr(out)))
    return out.reshape(D, m)


def dec_split(zmix: np.ndarray, k_dec: int):
    zmix = np.ascontiguousarray(zmix, np.int32).reshape(-1)
    m = len(zmix) // D
    outs = [np.zeros(D * m, np.int8) for _ in range(k_dec)]
    nz = np.zeros(k_dec, np.int32)
    _check(lib().ns_dec_split(_ptr(zmix), m, k_dec, _ptr_array(outs), _ptr(nz)))
    return [o.reshape(D, m) for o in outs], [bool(x) for x in nz]


class _LutInst(C.Structure):
    _fields_ = [("k", C.c_size_t), ("d", C.c_size_t), ("n_side", C.c_size_t), ("steps", C.c_size_t),
                ("ell", C.c_size_t), ("table", C.c_void_p)]


class _MemInst(C.Structure):
    _fields_ = [("k", C.c_size_t), ("d", C.c_size_t), ("n_side", C.c_size_t), ("steps", C.c_size_t),
                ("ell", C.c_size_t), ("init_sparse", C.c_int), ("ninit", C.c_size_t), ("init_addrs", C.c_void_p),
                ("init_vals", C.c_void_p)]


class _StepMem(C.Structure):
    _fields_ = [("n_lut", C.c_size_t), ("luts", C.c_void_p), ("n_mem", C.c_size_t), ("mems", C.c_void_p)]


class StepMemArray:
    """Per-step memory instances converted once to an ns_step_mem array (reusable across proofs)."""

    def __init__(self, mems):
        self.keep = []
        self.arr = _step_mem_array(mems, self.keep)
        self.n = len(mems)

    def __len__(self):
        return self.n


def _step_mem_array(mems, keep: list):
    """ns_step_mem[] for per-step memory instances (objects with .luts / .mems whose items carry
    k, d, n_side, steps, ell and .table (LUT) or .init (MEM: None = MemInit::Zero, else sorted
    (addr, value) pairs)); the arrays they point into are appended to `keep`."""
    arr = (_StepMem * len(mems))()
    for s, sm in enumerate(mems):
        luts = (_LutInst * max(1, len(sm.luts)))()
        for i, lt in enumerate(sm.luts):
            tab = np.ascontiguousarray(lt.table, np.uint64)
            keep.append(tab)
            luts[i] = _LutInst(lt.k, lt.d, lt.n_side, lt.steps, lt.ell, tab.ctypes.data)
        ms = (_MemInst * max(1, len(sm.mems)))()
        for i, mi in enumerate(sm.mems):
            if mi.init is None:
                ms[i] = _MemInst(mi.k, mi.d, mi.n_side, mi.steps, mi.ell, 0, 0, None, None)
            else:
                a = np.array([x for x, _ in mi.init] or [0], np.uint64)
                v = np.array([y for _, y in mi.init] or [0], np.uint64)
                keep += [a, v]
                ms[i] = _MemInst(mi.k, mi.d, mi.n_side, mi.steps, mi.ell, 1, len(mi.init), a.ctypes.data,
                                 v.ctypes.data)
        keep += [luts, ms]
        arr[s] = _StepMem(len(sm.luts), C.cast(luts, C.c_void_p), len(sm.mems), C.cast(ms, C.c_void_p))
    keep.append(arr)
    return arr


def set_shared_rounds(on: bool) -> None:
    """Shared launches for the fused rounds of concurrent provers on this GPU (process-wide)."""
    global _shared_on
    _check(lib().ns_set_shared_rounds(1 if on else 0))
    _shared_on = bool(on)


_shared_on = False


def shared_rounds_enabled() -> bool:
    """Whether the last set_shared_rounds call of this process turned shared rounds on."""
    return _shared_on


def shared_round_stats():
    """(launches issued, rounds carried) by the shared-round launcher since the process started."""
    a, b = C.c_uint64(0), C.c_uint64(0)
    _check(lib().ns_shared_round_stats(C.byref(a), C.byref(b)))
    return a.value, b.value


def ajtai_force_variant(variant: int) -> None:
    """Commitment GEMM tile shape: 0 timed choice (default), 1 16x16x64 tiles, 2 32x32x32 tiles on int8
    digits, 3 32x32x32 tiles on 2-bit packed digits."""
    _check(lib().ns_ajtai_force_variant(int(variant)))


class _Me(C.Structure):
    _fields_ = [("c", C.c_uint64 * (KAPPA * D)), ("m_in", C.c_size_t), ("X", C.c_void_p), ("ell_n", C.c_size_t),
                ("r", C.c_void_p), ("t", C.c_size_t), ("y", C.c_void_p), ("y_scalars", C.c_void_p),
                ("fold_digest", C.c_uint8 * 32)]


def _me_c(me, keep: list):
    """ns_me for a proof.MeInstance (input), its arrays kept alive in `keep`."""
    X = np.ascontiguousarray(np.asarray(me.X, np.uint64).reshape(-1))
    r = np.array([w for k in me.r for w in k] or [0], np.uint64)
    y = np.array([w for row in me.y for k in row for w in k] or [0], np.uint64)
    ys = np.array([w for k in me.y_scalars for w in k] or [0], np.uint64)
    keep += [X, r, y, ys]
    e = _Me()
    e.c[:] = [int(v) for v in me.c]
    e.m_in, e.X, e.ell_n, e.r = me.m_in, X.ctypes.data, len(me.r), r.ctypes.data
    e.t, e.y, e.y_scalars = len(me.y), y.ctypes.data, ys.ctypes.data
    e.fold_digest[:] = list(me.fold_digest)
    return e


def _me_out(m_in: int, ell_n: int, t: int, keep: list):
    X, r = np.zeros(max(1, D * m_in), np.uint64), np.zeros(max(1, 2 * ell_n), np.uint64)
    y, ys = np.zeros(max(1, 2 * 64 * t), np.uint64), np.zeros(max(1, 2 * t), np.uint64)
    keep.append((X, r, y, ys))
    e = _Me()
    e.m_in, e.X, e.ell_n, e.r, e.t, e.y, e.y_scalars = m_in, X.ctypes.data, ell_n, r.ctypes.data, t, y.ctypes.data, \
        ys.ctypes.data
    return e, (X, r, y, ys)


def _me_py(e, arrs):
    from .proof import MeInstance
    X, r, y, ys = arrs
    k = lambda a, i: (int(a[2 * i]), int(a[2 * i + 1]))  # noqa: E731
    return MeInstance(np.array(e.c[:], np.uint64), e.m_in, X[:D * e.m_in].reshape(D, e.m_in),
                      [k(r, i) for i in range(e.ell_n)], [[k(y, j * 64 + q) for q in range(64)] for j in range(e.t)],
                      [k(ys, j) for j in range(e.t)], bytes(e.fold_digest))


def rlc(session: "FoldingSession", mes, rho_coeffs, z_bufs, zmix_buf: "DeviceBuffer"):
    """ns_rlc: Pi_RLC of ME instances (proof.MeInstance) with rotation vectors rho_coeffs (k x 54 in
    [-2, 2]) over device digit planes; Z_mix lands in zmix_buf.  Returns the parent."""
    keep = []
    arr = (_Me * len(mes))(*[_me_c(m, keep) for m in mes])
    a = np.ascontiguousarray(np.asarray(rho_coeffs, np.int8).reshape(-1))
    out, arrs = _me_out(mes[0].m_in, len(mes[0].r), len(mes[0].y), keep)
    zp = (C.c_void_p * len(z_bufs))(*[b.data_ptr() for b in z_bufs])
    _check(lib().ns_rlc(session.h, arr, len(mes), a.ctypes.data, zp, C.byref(out), zmix_buf.data_ptr()))
    return _me_py(out, arrs)


def dec(session: "FoldingSession", parent, zmix_buf: "DeviceBuffer", k_dec: int, plane_bufs, bus_chunk: int = 0):
    """ns_dec: the k_dec children of a parent from Z_mix (planes written to plane_bufs).  Returns
    (children, (ok_y, ok_X, ok_c))."""
    keep = []
    pe = _me_c(parent, keep)
    outs = [_me_out(parent.m_in, len(parent.r), len(parent.y), keep) for _ in range(k_dec)]
    arr = (_Me * k_dec)(*[o[0] for o in outs])
    pp = (C.c_void_p * k_dec)(*[b.data_ptr() for b in plane_bufs])
    ok = (C.c_int * 3)()
    _check(lib().ns_dec(session.h, C.byref(pe), zmix_buf.data_ptr(), k_dec, pp, bus_chunk, arr, ok))
    return [_me_py(arr[i], outs[i][1]) for i in range(k_dec)], tuple(bool(x) for x in ok)


class ProgramIO:
    """ProgramIO (neo-memory/src/output_check.rs:92-171): claimed (address, value) pairs, kept in
    address order; a second claim at an address replaces the first (with_claim), or raises
    (try_with_claim)."""

    def __init__(self):
        self.claims: dict[int, int] = {}

    def with_claim(self, addr: int, value: int) -> "ProgramIO":
        self.claims[int(addr)] = int(value) % Q
        return self

    with_output = with_claim
    with_input = with_claim

    def try_with_claim(self, addr: int, value: int) -> "ProgramIO":
        if int(addr) in self.claims:
            raise ValueError(f"Duplicate address {addr}")
        return self.with_claim(addr, value)

    def sorted_claims(self):
        return sorted(self.claims.items())


class OutputBindingConfig:
    """OutputBindingConfig (neo-fold/src/output_binding.rs:23-60): the address width, the claims
    and the index of the bound Twist instance among the last step's memory instances."""

    def __init__(self, num_bits: int, program_io: ProgramIO, mem_idx: int = 0):
        self.num_bits, self.program_io, self.mem_idx = num_bits, program_io, mem_idx

    def with_mem_idx(self, mem_idx: int) -> "OutputBindingConfig":
        self.mem_idx = mem_idx
        return self


class _ObC(C.Structure):
    _fields_ = [("mem_idx", C.c_size_t), ("num_bits", C.c_size_t), ("n_claims", C.c_size_t),
                ("addrs", C.c_void_p), ("values", C.c_void_p), ("final_state", C.c_void_p)]


def _ob_c(cfg: OutputBindingConfig, final_state, keep: list) -> _ObC:
    cl = cfg.program_io.sorted_claims()
    a = np.array([x for x, _ in cl] or [0], np.uint64)
    v = np.array([y for _, y in cl] or [0], np.uint64)
    keep += [a, v]
    fs = None
    if final_state is not None:
        fs = np.ascontiguousarray(np.asarray(final_state, np.uint64))
        if fs.size != 1 << cfg.num_bits:
            raise NsError(-1, f"final_memory_state.len()={fs.size} != 2^num_bits={1 << cfg.num_bits}")
        keep.append(fs)
    o = _ObC(cfg.mem_idx, cfg.num_bits, len(cl), a.ctypes.data, v.ctypes.data, None if fs is None else fs.ctypes.data)
    keep.append(o)
    return o


def shard_verify(ccs: "Ccs", xs, proof: bytes, mem=None, mcs_commitments=None, m_in: int = 1,
                 ob: "OutputBindingConfig | None" = None) -> None:
    """ns_shard_verify (ns_shard_verify_ob when ob is given): replay the transcript of a canonical
    shard proof and check it (host only).  xs: each step's public input (m_in values); mem: per-step
    memory instances or None; ob: the output binding the proof must carry.  Raises NsError with the
    reason when the proof is rejected."""
    xs = [np.ascontiguousarray(np.asarray(x, np.uint64)[:m_in]) for x in xs]
    keep = []
    marr = _step_mem_array(mem, keep) if mem is not None else None
    mc = None
    if mcs_commitments is not None:
        mc = np.ascontiguousarray(np.asarray(mcs_commitments, np.uint64).reshape(-1))
    buf = (C.c_uint8 * len(proof)).from_buffer_copy(proof)
    mp = None if marr is None else C.cast(marr, C.c_void_p)
    mcp = None if mc is None else mc.ctypes.data
    if ob is None:
        _check(lib().ns_shard_verify(ccs.h, _ptr_array(xs), len(xs), m_in, mp, mcp, buf, len(proof)))
    else:
        obc = _ob_c(ob, None, keep)
        _check(lib().ns_shard_verify_ob(ccs.h, _ptr_array(xs), len(xs), m_in, mp, C.byref(obc), mcp, buf,
                                        len(proof)))


class FoldingSession:
    """Mirror of neo-fold FoldingSession (session.rs:668-1953).

    new_ajtai_seeded(ccs, seed) -> add_step_io(x, w[, mem]) per step -> fold_and_prove() returns
    the canonical proof encoding; the accumulator and witnesses stay in HBM during the fold.  A
    step's `mem` carries its Twist/Shout instances over the shared CPU bus (Route A); every step
    of a shard either has one or none does.
    """

    def __init__(self, ccs: Ccs, pp: PublicParams):
        self.ccs, self.pp = ccs, pp
        h = C.c_void_p()
        _check(lib().ns_prover_new(ccs.h, pp.h, C.byref(h)))
        self.h = h
        self.steps: list[np.ndarray] = []
        self.mem: list = []
        self.m_in: int | None = None
        self.last_seconds = 0.0

    @classmethod
    def new_ajtai_seeded(cls, inst, seed: bytes):
        ccs = Ccs(inst)
        pp = PublicParams(inst.m, seed=seed)
        return cls(ccs, pp)

    def add_step_io(self, public_input, witness, mem=None):
        x = np.asarray(public_input, np.uint64)
        w = np.asarray(witness, np.uint64)
        if self.m_in is None:
            self.m_in = len(x)
        elif self.m_in != len(x):
            raise NsError(-1, "all steps must share the same m_in")
        z = np.concatenate([x, w]).astype(np.uint64)
        if len(z) != self.ccs.m:
            raise NsError(-1, f"len(x) + len(w) = {len(z)} but CCS.m = {self.ccs.m}")
        self.steps.append(np.ascontiguousarray(z))
        self.mem.append(mem)

    def add_step_z(self, z, m_in=1, mem=None):
        self.add_step_io(z[:m_in], z[m_in:], mem)

    def fold_and_prove(self) -> bytes:
        return self._prove_host(None, None)

    def fold_and_prove_with_output_binding(self, ob_cfg: OutputBindingConfig, final_memory_state) -> bytes:
        """fold_shard_prove_with_output_binding (shard.rs:2352-2385): the last step also proves that
        Twist instance ob_cfg.mem_idx ends in final_memory_state (2^num_bits values) and that the
        claimed outputs match it."""
        if not self.mem or any(x is None for x in self.mem):
            raise NsError(-1, "output binding requires memory instances on every step")
        return self._prove_host(ob_cfg, final_memory_state)

    def _prove_host(self, ob_cfg, final_state) -> bytes:
        out = C.POINTER(C.c_uint8)()
        n = C.c_size_t()
        sec = C.c_double()
        if ob_cfg is not None:
            keep = []
            arr = _step_mem_array(self.mem, keep)
            obc = _ob_c(ob_cfg, final_state, keep)
            _check(lib().ns_prove_shard_ob(self.h, _ptr_array(self.steps), 0, len(self.steps), self.m_in or 0,
                                           C.cast(arr, C.c_void_p), C.byref(obc), C.byref(out), C.byref(n),
                                           C.byref(sec)))
        elif any(x is not None for x in self.mem):
            if any(x is None for x in self.mem):
                raise NsError(-1, "every step needs its memory instances once one step has them")
            keep = []
            arr = _step_mem_array(self.mem, keep)
            _check(lib().ns_prove_shard_mem(self.h, _ptr_array(self.steps), 0, len(self.steps), self.m_in or 0,
                                            C.cast(arr, C.c_void_p), C.byref(out), C.byref(n), C.byref(sec)))
        else:
            _check(lib().ns_prove_shard(self.h, _ptr_array(self.steps), len(self.steps), self.m_in or 0,
                                        C.byref(out), C.byref(n), C.byref(sec)))
        data = bytes(C.cast(out, C.POINTER(C.c_uint8 * n.value)).contents)
        lib().ns_free(C.cast(out, C.c_void_p))
        self.last_seconds = sec.value
        return data

    def prove_device(self, z_ptrs, m_in: int = 1, mem=None, ob=None, final_state=None) -> bytes:
        """fold_and_prove over step witnesses already resident in HBM (device pointers to u64[m]);
        mem: per-step memory instances (Route A; a list or a StepMemArray covering at least the
        proved steps) or None; ob / final_state: output binding on the last step, or None."""
        arr = (C.c_void_p * len(z_ptrs))(*z_ptrs)
        out = C.POINTER(C.c_uint8)()
        n = C.c_size_t()
        sec = C.c_double()
        if ob is not None:
            if mem is None:
                raise NsError(-1, "output binding requires memory instances")
            sm = mem if isinstance(mem, StepMemArray) else StepMemArray(mem)
            if len(sm) < len(z_ptrs):
                raise ValueError("fewer memory steps than witnesses")
            keep = []
            obc = _ob_c(ob, final_state, keep)
            _check(lib().ns_prove_shard_ob(self.h, arr, 1, len(z_ptrs), m_in, C.cast(sm.arr, C.c_void_p),
                                           C.byref(obc), C.byref(out), C.byref(n), C.byref(sec)))
        elif mem is not None:
            sm = mem if isinstance(mem, StepMemArray) else StepMemArray(mem)
            if len(sm) < len(z_ptrs):
                raise ValueError("fewer memory steps than witnesses")
            _check(lib().ns_prove_shard_mem(self.h, arr, 1, len(z_ptrs), m_in, C.cast(sm.arr, C.c_void_p),
                                            C.byref(out), C.byref(n), C.byref(sec)))
        else:
            _check(lib().ns_prove_shard_device(self.h, arr, len(z_ptrs), m_in, C.byref(out), C.byref(n),
                                               C.byref(sec)))
        data = bytes(C.cast(out, C.POINTER(C.c_uint8 * n.value)).contents)
        lib().ns_free(C.cast(out, C.c_void_p))
        self.last_seconds = sec.value
        return data

    def rebind(self, ccs: "Ccs", pp: "PublicParams"):
        """Reuse this session's device buffers for another CCS / PP (same semantics as a new session)."""
        _check(lib().ns_prover_rebind(self.h, ccs.h, pp.h))
        self.ccs, self.pp = ccs, pp
        self.steps, self.mem, self.m_in = [], [], None

    def set_timing(self, on: bool):
        _check(lib().ns_prover_set_timing(self.h, int(on)))

    def kernel_stats(self):
        ms = np.zeros(9, np.float64)
        cnt = np.zeros(9, np.uint64)
        _check(lib().ns_prover_kernel_stats(self.h, _ptr(ms), _ptr(cnt), 9))
        return ms.tolist(), [int(c) for c in cnt]

    PHASES = ["bind_inputs", "oracle_build", "row_rounds", "y_products", "ajtai_rounds", "me_outputs",
              "rlc_bind", "zmix", "dec", "encode", "intake", "acc_swap"]

    def phase_stats(self) -> dict:
        out = np.zeros(12, np.float64)
        _check(lib().ns_prover_phase_stats(self.h, _ptr(out), 12))
        return {k: round(float(v), 5) for k, v in zip(self.PHASES, out)}

    def phase_perms(self) -> dict:
        """Poseidon2 permutations of the host transcript per phase of the last prove."""
        out = np.zeros(12, np.uint64)
        _check(lib().ns_prover_phase_perms(self.h, _ptr(out), 12))
        return {k: int(v) for k, v in zip(self.PHASES, out)}

    def __del__(self):
        if getattr(self, "h", None):
            lib().ns_prover_free(self.h)
            self.h = None
