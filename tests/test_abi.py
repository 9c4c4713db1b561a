"""The C-ABI library loads and exports every entry point include/graft_heff.h declares (CPU only)."""
import ctypes
import os
import re
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "graft_heff.h")


def declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(graft_[a-z0-9_]+)\s*\(", text)))


def test_header_declares_entry_points():
    names = declared_functions()
    for required in ["graft_plan_create", "graft_plan_execute", "graft_heff_apply",
                     "graft_env_update", "graft_contract_pair", "graft_vec_dotc"]:
        assert required in names


def test_library_exports_every_declared_symbol():
    from graft_amd import _native as nat
    if not os.path.exists(nat.LIB_PATH):
        pytest.skip("library not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(nat.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    # and they are exported with C linkage (no mangling)
    out = subprocess.run(["nm", "-D", "--defined-only", nat.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(l.split()[-1] for l in out.splitlines() if l.strip())
    assert set(declared_functions()) <= exported


def test_binding_loads_and_reports_errors_without_gpu():
    from graft_amd import _native as nat
    if not os.path.exists(nat.LIB_PATH):
        pytest.skip("library not built")
    L = nat.lib()
    assert b"gfx950" in L.graft_version()
    # a null context is rejected with a message, without touching a device
    assert L.graft_ctx_set_stream(None, None) == 1
    assert b"null" in L.graft_last_error()
    assert L.graft_num_tile_configs() >= 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py runs on one GPU: --gpus N with N > 1 prints one JSON line saying
    so and exits with status 2, before any device work."""
    import json
    import sys
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2, (r.returncode, r.stderr[-400:])
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
