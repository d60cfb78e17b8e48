"""The C-ABI library loads (no GPU needed) and exports every symbol that
include/pdes.h declares."""
import ctypes as C
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "pdes.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(pdes_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from pdes_amd import _lib
    assert os.path.exists(_lib.LIB_PATH), "build the library first (__graft_entry__.build())"
    L = C.CDLL(_lib.LIB_PATH)
    missing = [s for s in declared_symbols() if not hasattr(L, s)]
    assert not missing, f"missing exports: {missing}"
    assert len(declared_symbols()) >= 20


def test_binding_covers_header():
    from pdes_amd import _lib
    declared = set(declared_symbols())
    bound = set(_lib.exported_symbols())
    assert declared <= bound, f"unbound: {sorted(declared - bound)}"


def test_package_default_hardware_queue_count():
    """Importing the package gives HIP 16 hardware queues where the
    environment sets none (a batch runs a stream per operator group) and
    leaves an explicit setting alone; run in child processes, so this
    process's environment is untouched."""
    import subprocess
    import sys
    code = "import os; from pdes_amd import _lib; print(os.environ['GPU_MAX_HW_QUEUES'])"
    pkg = os.path.join(ROOT, "pdesolver.jl_amd")
    for given, expected in ((None, "16"), ("4", "4")):
        env = {k: v for k, v in os.environ.items() if k != "GPU_MAX_HW_QUEUES"}
        if given is not None:
            env["GPU_MAX_HW_QUEUES"] = given
        env["PYTHONPATH"] = pkg + os.pathsep + env.get("PYTHONPATH", "")
        r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env, timeout=120)
        assert r.returncode == 0, r.stderr
        assert r.stdout.strip() == expected
