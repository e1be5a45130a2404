"""The MI355X library loads on a host without a GPU and exports every entry point of the C header.

No compute call is made here; the GPU tests drive the entry points.
"""
import ctypes as C
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "nightstream.h")
LIB = os.path.join(ROOT, "nightstream_amd", "libnightstream.so")


def header_symbols():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ns_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    syms = header_symbols()
    for required in ("ns_ajtai_commit_batch", "ns_oracle_evals", "ns_oracle_fold", "ns_prove_shard",
                     "ns_rlc_zmix", "ns_dec_split", "ns_pp_setup_seeded"):
        assert required in syms


@pytest.mark.skipif(not os.path.exists(LIB), reason="libnightstream.so not built (run __graft_entry__.build())")
def test_library_exports_every_header_symbol():
    lib = C.CDLL(LIB)
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, f"missing exports: {missing}"


def test_python_wrapper_binds_header_symbols():
    import nightstream_amd
    assert set(nightstream_amd.EXPORTS) <= set(header_symbols())
    assert set(header_symbols()) <= set(nightstream_amd.EXPORTS)
