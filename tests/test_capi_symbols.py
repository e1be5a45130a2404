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


@pytest.mark.skipif(not os.path.exists(LIB), reason="libnightstream.so not built (run __graft_entry__.build())")
def test_library_loads_once_under_threads():
    """16 threads loading and binding the library at once (as a caller building oracles on a
    thread pool does) all get the one library object, with the Twist/Shout signatures set."""
    import subprocess
    import sys
    code = (
        "import threading, ctypes\n"
        "import nightstream_amd as ns\n"
        "from nightstream_amd import twist\n"
        "got = []\n"
        "barrier = threading.Barrier(16)\n"
        "def go():\n"
        "    barrier.wait()\n"
        "    got.append(twist._bind())\n"
        "ts = [threading.Thread(target=go) for _ in range(16)]\n"
        "[t.start() for t in ts]\n"
        "[t.join() for t in ts]\n"
        "assert len(got) == 16 and all(g is got[0] for g in got) and got[0] is ns.lib()\n"
        "assert got[0].ns_ts_batch_round.argtypes is not None\n"
        "assert got[0].ns_ts_num_rounds.restype is ctypes.c_size_t\n"
        "print('ok')\n")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0 and r.stdout.strip() == "ok", r.stderr
