"""NeoParams auto-lambda for R1CS (neo-params/src/lib.rs:186-304) against exact Python integers.

With s = 2 the search keeps the largest lambda <= 127 (and >= 96) whose slack
floor(log2(q^2 / (2^lambda * ell * d_sc))) is at least 2, where ell = ceil(log2(54 * n_pad)) and
d_sc = 1 + max(2, 2b, 2) = 5.  Worked by hand: lambda = 119 for n in [2^10, 2^18], 118 at 2^20.
"""
import ctypes as C

import pytest

from kmath import Q


class Params(C.Structure):
    _fields_ = [("q", C.c_uint64), ("eta", C.c_uint32), ("d", C.c_uint32), ("kappa", C.c_uint32),
                ("m", C.c_uint64), ("b", C.c_uint32), ("k_rho", C.c_uint32), ("B", C.c_uint64),
                ("T", C.c_uint32), ("s", C.c_uint32), ("lam", C.c_uint32)]


def _py_lambda(n_rows):
    n_pad = 2 if n_rows == 0 else max(2, 1 << (n_rows - 1).bit_length())
    ell = (54 * n_pad - 1).bit_length()
    ld = ell * 5
    for lam in range(127, 95, -1):
        rhs = (1 << lam) * ld
        if Q * Q < rhs:
            continue
        slack = 0
        while Q * Q >= rhs << (slack + 1):
            slack += 1
        if slack >= 2:
            return lam
    return None


@pytest.mark.parametrize("n", [0, 1, 5, 1 << 10, (1 << 10) + 1, 1 << 12, 1 << 16, 1 << 18, 1 << 20, 1 << 24])
def test_auto_lambda(oracle, n):
    p = Params()
    L = oracle.lib()
    L.or_params_auto_r1cs.argtypes = [C.c_size_t, C.POINTER(Params)]
    rc = L.or_params_auto_r1cs(n, C.byref(p))
    assert rc == 0
    assert p.lam == _py_lambda(n)
    assert (p.q, p.d, p.kappa, p.b, p.k_rho, p.B, p.T, p.s) == (Q, 54, 16, 2, 12, 4096, 216, 2)


def test_lambda_at_the_measured_sizes(oracle):
    for e in range(10, 19):
        assert _py_lambda(1 << e) == 119
    assert _py_lambda(1 << 20) == 118


@pytest.mark.parametrize("lam,ell", [(119, 16), (120, 16), (121, 16), (127, 16), (100, 30), (124, 2), (60, 16)])
def test_extension_check_slack(oracle, lam, ell):
    L = oracle.lib()
    L.or_extension_check.argtypes = [C.POINTER(Params), C.c_uint32, C.c_uint32, C.POINTER(C.c_uint32),
                                     C.POINTER(C.c_int32)]
    p = Params(q=Q, d=54, b=2, lam=lam)
    sm, sl = C.c_uint32(), C.c_int32()
    rc = L.or_extension_check(C.byref(p), ell, 5, C.byref(sm), C.byref(sl))
    rhs = (1 << lam) * ell * 5
    exp_smin = 1 if Q >= rhs else (2 if Q * Q >= rhs else 3)
    if exp_smin > 2:
        assert rc != 0
        return
    assert rc == 0 and sm.value == exp_smin
    exp_slack = -1 if Q * Q < rhs else max(s for s in range(0, 130) if Q * Q >= rhs << s)
    assert sl.value == exp_slack
