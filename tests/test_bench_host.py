"""Host-side logic of bench.py (no GPU): the one-GPU refusal, the hardware
queue count it gives the process, the measured traffic it reads from the
committed counter summary, and the algorithmic bytes and flops its roofline
divides by, checked against counts of the 256 x 256 grid's elements and
edges."""
import importlib.util
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _load_bench(monkeypatch, queues):
    """bench.py as a fresh module, with GPU_MAX_HW_QUEUES as given (None:
    unset); monkeypatch restores the variable afterwards."""
    if queues is None:
        monkeypatch.delenv("GPU_MAX_HW_QUEUES", raising=False)
    else:
        monkeypatch.setenv("GPU_MAX_HW_QUEUES", queues)
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


@pytest.mark.parametrize("given,expected", [(None, "16"), ("4", "16"), ("16", "16"), ("32", "32")])
def test_hardware_queue_count(monkeypatch, given, expected):
    """The batched and served solves run a stream per operator group; bench.py
    raises the process's hardware queue count to at least 16 and leaves a
    larger setting alone."""
    _load_bench(monkeypatch, given)
    assert os.environ["GPU_MAX_HW_QUEUES"] == expected


def test_traffic_is_read_from_the_committed_counter_summary(monkeypatch):
    b = _load_bench(monkeypatch, None)
    rows = {}
    with open(os.path.join(ROOT, "profiles", "euler_residual_256x256_hbm_traffic.md")) as f:
        for line in f:
            for key in ("p2_roe", "p4_roe", "p2_ir", "p4_ir"):
                if line.startswith(f"| {key} |"):
                    rows[key] = float(line.strip().strip("|").split("|")[-1])
    assert set(rows) == {"p2_roe", "p4_roe", "p2_ir", "p4_ir"}
    for key, val in rows.items():
        p, flux = int(key[1]), key.split("_")[1]
        assert b.traffic_for(p, flux) == val
    assert b.traffic_for(3, "roe") is None


@pytest.mark.parametrize("p,flux", [(2, "roe"), (4, "ir")])
def test_algorithmic_bytes_and_flops_from_grid_counts(monkeypatch, p, flux):
    """The 256 x 256 quarter annulus has 2 n^2 triangles, 2 n (n - 1) + n^2
    interior edges and 4 n boundary edges; the diagonal-E triangles have 12
    (p = 2) and 27 (p = 4) volume nodes and p + 2 nodes per face.  Bytes: q
    read and res written (4 unknowns per node), one
    2 x 2 metric per element (straight sides), 16 B per element face record,
    one normal per interface face node, a normal and a coordinate per boundary
    face node.  Flops: 73 per node pair and 18 per node in the volume, 207 per
    face node of each edge."""
    b = _load_bench(monkeypatch, None)
    m, op, ph, q = b.make(p, flux)
    n = 256
    nel, n_int, n_bnd = 2 * n * n, 2 * n * (n - 1) + n * n, 4 * n
    npe = {2: 12, 4: 27}[p]
    nf = p + 2
    assert (m.numEl, m.numInterfaces, m.numBoundaryFaces) == (nel, n_int, n_bnd)
    assert (op.numnodes, op.face.numnodes) == (npe, nf)
    per_el = 2 * 4 * npe * 8 + 2 * 2 * 8 + 3 * 16
    faces = n_int * nf * 2 * 8 + n_bnd * nf * 2 * 2 * 8
    assert b.residual_bytes_per_launch(m, op) == per_el * nel + faces
    assert b.residual_bytes_per_launch(m, op) == {2: 123830272, 4: 256000000}[p]
    if flux == "ir":
        pairs = npe * (npe - 1) // 2
        flops = (pairs * 73 + npe * 18) * nel + nf * (n_int + n_bnd) * 207
        assert b.residual_flops_per_launch(m, op) == flops == 3666981888
