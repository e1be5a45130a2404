"""Twist/Shout round oracles on the GPU (ctypes mirror of the ns_ts_* C ABI).

Each TsOracle implements the RoundOracle contract of neo-reductions/src/sumcheck.rs:17-34
(evals / fold / num_rounds / degree_bound).  Sparse time columns are passed as
(indices, [K values]) pairs, the SparseIdxVec of neo-memory/src/sparse_time.rs; K values are
(c0, c1) tuples of canonical Goldilocks integers or, passed without conversion, an (n, 2) uint64
array.  The constructors take the arguments of the
reference constructors (neo-memory/src/twist_oracle.rs) in the same order.
"""
from __future__ import annotations

import ctypes as C
import threading

import numpy as np

from . import NsError, _check, lib


class SparseCol(C.Structure):
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

    def col(self, c):
        idx = np.ascontiguousarray(np.asarray(c[0], np.uint64).reshape(-1)) if len(c[0]) else np.zeros(1, np.uint64)
        vals = _kbuf(c[1])
        self.items += [idx, vals]
        return SparseCol(len(c[0]), idx.ctypes.data_as(C.POINTER(C.c_uint64)),
                         vals.ctypes.data_as(C.POINTER(C.c_uint64)))

    def cols(self, cs):
        arr = (SparseCol * max(1, len(cs)))(*[self.col(c) for c in cs])
        self.items.append(arr)
        return arr

    def ks(self, ks):
        a = _kbuf(ks)
        self.items.append(a)
        return a.ctypes.data_as(C.c_void_p)


_BOUND = False
_BIND_LOCK = threading.Lock()


def _bind():
    """The library with the ns_ts_* signatures set (once, under a lock: oracles may be built on
    several threads)."""
    L = lib()
    if _BOUND:
        return L
    with _BIND_LOCK:
        if not _BOUND:
            _set_argtypes(L)
    return L


def _set_argtypes(L):
    global _BOUND
    vp, sz, S = C.c_void_p, C.c_size_t, SparseCol
    out = C.POINTER(vp)
    L.ns_ts_product_new.argtypes = [vp, sz, sz, sz, out]
    L.ns_ts_shout_value_new.argtypes = [vp, sz, S, S, vp, out]
    L.ns_ts_index_adapter_new.argtypes = [vp, sz, S, vp, sz, vp, vp, out]
    L.ns_ts_bitness_new.argtypes = [vp, sz, vp, vp, sz, out]
    L.ns_ts_twist_read_new.argtypes = [vp, sz, S, S, vp, S, S, vp, sz, vp, vp, out]
    L.ns_ts_twist_write_new.argtypes = [vp, sz, S, S, S, vp, sz, vp, vp, out]
    L.ns_ts_val_eval_new.argtypes = [sz, vp, sz, S, S, vp, vp, vp, out]
    L.ns_ts_total_inc_new.argtypes = [sz, vp, sz, S, S, vp, vp, out]
    L.ns_ts_read_addr_new.argtypes = [vp, vp, sz, vp, sz, S, S, vp, S, vp, S, sz, out]
    L.ns_ts_write_addr_new.argtypes = [vp, vp, sz, vp, sz, S, S, vp, S, sz, out]
    L.ns_ts_addr_lookup_new.argtypes = [vp, sz, S, vp, sz, vp, sz, vp, out]
    L.ns_ts_evals.argtypes = [vp, vp, sz, vp]
    L.ns_ts_fold.argtypes = [vp, vp]
    L.ns_ts_batch_round.argtypes = [vp, sz, vp]
    L.ns_ts_batch_fold.argtypes = [vp, sz, vp]
    L.ns_ts_num_rounds.argtypes = [vp]
    L.ns_ts_num_rounds.restype = sz
    L.ns_ts_degree_bound.argtypes = [vp]
    L.ns_ts_degree_bound.restype = sz
    L.ns_ts_free.argtypes = [vp]
    L.ns_ts_free.restype = None
    _BOUND = True


class TsOracle:
    """Device-resident Twist/Shout oracle; `kind` names the reference oracle (see module doc)."""

    def __init__(self, kind, *args):
        L = _bind()
        k = _Keep()
        self._keep = k
        h = C.c_void_p()
        claim = np.zeros(2, np.uint64)
        cp = claim.ctypes.data_as(C.c_void_p)
        if kind == "product":
            factors, degree_bound = args
            bufs = [_kbuf(f) for f in factors]
            k.items += bufs
            ptrs = (C.c_void_p * len(bufs))(*[b.ctypes.data for b in bufs])
            k.items.append(ptrs)
            rc = L.ns_ts_product_new(ptrs, len(factors), len(factors[0]), degree_bound, C.byref(h))
        elif kind == "shout_value":
            r_cycle, has_lookup, val = args
            rc = L.ns_ts_shout_value_new(k.ks(r_cycle), len(r_cycle), k.col(has_lookup), k.col(val), cp, C.byref(h))
        elif kind == "index_adapter":
            r_cycle, has_lookup, addr_bits, r_addr = args
            rc = L.ns_ts_index_adapter_new(k.ks(r_cycle), len(r_cycle), k.col(has_lookup), k.cols(addr_bits),
                                           len(addr_bits), k.ks(r_addr), cp, C.byref(h))
        elif kind == "bitness":
            r_cycle, cols, weights = args
            rc = L.ns_ts_bitness_new(k.ks(r_cycle), len(r_cycle), k.cols(cols), k.ks(weights), len(cols), C.byref(h))
        elif kind == "twist_read":
            r_cycle, has_read, rv, ra_bits, has_write, inc, wa_bits, r_addr, init_at = args
            rc = L.ns_ts_twist_read_new(k.ks(r_cycle), len(r_cycle), k.col(has_read), k.col(rv), k.cols(ra_bits),
                                        k.col(has_write), k.col(inc), k.cols(wa_bits), len(r_addr), k.ks(r_addr),
                                        k.ks([init_at]), C.byref(h))
        elif kind == "twist_write":
            r_cycle, has_write, wv, inc, wa_bits, r_addr, init_at = args
            rc = L.ns_ts_twist_write_new(k.ks(r_cycle), len(r_cycle), k.col(has_write), k.col(wv), k.col(inc),
                                         k.cols(wa_bits), len(r_addr), k.ks(r_addr), k.ks([init_at]), C.byref(h))
        elif kind == "val_eval":
            ell_n, wa_bits, has_write, inc, r_addr, r_time = args
            rc = L.ns_ts_val_eval_new(ell_n, k.cols(wa_bits), len(wa_bits), k.col(has_write), k.col(inc),
                                      k.ks(r_addr), k.ks(r_time), cp, C.byref(h))
        elif kind == "total_inc":
            ell_n, wa_bits, has_write, inc, r_addr = args
            rc = L.ns_ts_total_inc_new(ell_n, k.cols(wa_bits), len(wa_bits), k.col(has_write), k.col(inc),
                                       k.ks(r_addr), cp, C.byref(h))
        elif kind in ("read_addr", "write_addr"):
            init = args[0]
            ia = np.array([a for a, _ in init] or [0], np.uint64)
            k.items.append(ia)
            iv = k.ks([v for _, v in init])
            if kind == "read_addr":
                _, r_cycle, has_read, rv, ra_bits, has_write, wa_bits, inc = args
                rc = L.ns_ts_read_addr_new(ia.ctypes.data, iv, len(init), k.ks(r_cycle), len(r_cycle), k.col(has_read),
                                           k.col(rv), k.cols(ra_bits), k.col(has_write), k.cols(wa_bits), k.col(inc),
                                           len(ra_bits), C.byref(h))
            else:
                _, r_cycle, has_write, wv, wa_bits, inc = args
                rc = L.ns_ts_write_addr_new(ia.ctypes.data, iv, len(init), k.ks(r_cycle), len(r_cycle),
                                            k.col(has_write), k.col(wv), k.cols(wa_bits), k.col(inc), len(wa_bits),
                                            C.byref(h))
        elif kind == "addr_lookup":
            addr_bits, has_lookup, table, r_cycle = args
            rc = L.ns_ts_addr_lookup_new(k.cols(addr_bits), len(addr_bits), k.col(has_lookup), k.ks(table), len(table),
                                         k.ks(r_cycle), len(r_cycle), cp, C.byref(h))
        else:
            raise NsError(-1, f"unknown oracle kind {kind}")
        _check(rc)
        self.h = h
        self.claim = (int(claim[0]), int(claim[1]))

    def num_rounds(self) -> int:
        return _bind().ns_ts_num_rounds(self.h)

    def degree_bound(self) -> int:
        return _bind().ns_ts_degree_bound(self.h)

    def evals(self, xs):
        xb = _kbuf(xs)
        out = np.zeros(2 * len(xs), np.uint64)
        _check(_bind().ns_ts_evals(self.h, xb.ctypes.data, len(xs), out.ctypes.data))
        o = [int(v) for v in out]
        return [(o[2 * i], o[2 * i + 1]) for i in range(len(xs))]

    def fold(self, r):
        rb = np.array([int(r[0]), int(r[1])], np.uint64)
        _check(_bind().ns_ts_fold(self.h, rb.ctypes.data))

    def __del__(self):
        try:
            if getattr(self, "h", None):
                _bind().ns_ts_free(self.h)
        except Exception:
            pass


def _handles(oracles):
    return (C.c_void_p * max(1, len(oracles)))(*[o.h.value for o in oracles])


def batch_round(oracles):
    """One round of the batched sum-check (neo-reductions/src/sumcheck.rs:390-407): for every
    oracle, its round polynomial's values at 0..degree_bound; all rounds are enqueued before
    the host waits.  Every oracle must have the same number of rounds left."""
    degs = [o.degree_bound() for o in oracles]
    out = np.zeros(2 * max(1, sum(d + 1 for d in degs)), np.uint64)
    _check(_bind().ns_ts_batch_round(_handles(oracles), len(oracles), out.ctypes.data))
    o, res, w = [int(v) for v in out], [], 0
    for d in degs:
        res.append([(o[2 * (w + i)], o[2 * (w + i) + 1]) for i in range(d + 1)])
        w += d + 1
    return res


def batch_fold(oracles, r):
    """Fold every oracle with the shared challenge r (sumcheck.rs:436-442)."""
    rb = np.array([int(r[0]), int(r[1])], np.uint64)
    _check(_bind().ns_ts_batch_fold(_handles(oracles), len(oracles), rb.ctypes.data))


# ---------------- shared CPU bus (neo-fold/src/memory_sidecar/cpu_bus.rs) ----------------
def bus_decode(Z) -> np.ndarray:
    """decode_vector: z[c] = sum_rho 2^rho Z[rho][c] for a [54][m] int8 digit matrix."""
    Z = np.ascontiguousarray(Z, np.int8)
    out = np.zeros(Z.shape[1], np.uint64)
    L = lib()
    L.ns_bus_decode.argtypes = [C.c_void_p, C.c_size_t, C.c_void_p]
    _check(L.ns_bus_decode(Z.ctypes.data, Z.shape[1], out.ctypes.data))
    return out


def bus_time_cols(Z, m_in, bus_base, chunk, bus_cols, steps_len, ell_n):
    """build_time_sparse_from_bus_col for every bus column: [(idx array, [K values])]."""
    Z = np.ascontiguousarray(Z, np.int8)
    idx = np.zeros(max(1, bus_cols * chunk), np.uint64)
    vals = np.zeros(max(2, 2 * bus_cols * chunk), np.uint64)
    nnz = np.zeros(max(1, bus_cols), np.uint64)
    L = lib()
    L.ns_bus_time_cols.argtypes = [C.c_void_p] + [C.c_size_t] * 7 + [C.c_void_p] * 3
    _check(L.ns_bus_time_cols(Z.ctypes.data, Z.shape[1], m_in, bus_base, chunk, bus_cols, steps_len, ell_n,
                              idx.ctypes.data, vals.ctypes.data, nnz.ctypes.data))
    out = []
    for c in range(bus_cols):
        n = int(nnz[c])
        base = c * chunk
        out.append((idx[base:base + n].copy(),
                    [(int(vals[2 * (base + i)]), int(vals[2 * (base + i) + 1])) for i in range(n)]))
    return out


def bus_openings(Z, m_in, bus_base, chunk, bus_cols, r):
    """append_bus_openings_to_me_instance: (y rows padded to 64, y_scalars) per bus column."""
    Z = np.ascontiguousarray(Z, np.int8)
    rb = _kbuf(r)
    y = np.zeros(max(2, bus_cols * 128), np.uint64)
    ys = np.zeros(max(2, bus_cols * 2), np.uint64)
    L = lib()
    L.ns_bus_openings.argtypes = [C.c_void_p] + [C.c_size_t] * 5 + [C.c_void_p, C.c_size_t, C.c_void_p, C.c_void_p]
    _check(L.ns_bus_openings(Z.ctypes.data, Z.shape[1], m_in, bus_base, chunk, bus_cols, rb.ctypes.data, len(r),
                             y.ctypes.data, ys.ctypes.data))
    yl, ysl = [int(v) for v in y], [int(v) for v in ys]
    return ([[(yl[2 * (c * 64 + i)], yl[2 * (c * 64 + i) + 1]) for i in range(64)] for c in range(bus_cols)],
            [(ysl[2 * c], ysl[2 * c + 1]) for c in range(bus_cols)])
