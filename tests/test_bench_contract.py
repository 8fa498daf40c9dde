"""The bench.py contract pieces that run without a GPU."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_exits_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_sizes_and_work_of_the_h2o32_step():
    """N, o, v, naux of (H2O)32 in def2-SVP / def2-SVP-RI and the algorithmic
    work: W_E = 2 [o(o+1)/2] v^2 naux, W_B = 2 naux o N^2 + 2 naux o v N + 2 naux^2 o v
    with a dense metric, and naux^2 o v for the metric product with the
    triangular Cholesky factor that the step applies."""
    N, o, v, naux = bench.sizes(32)
    assert (N, o, v, naux) == (768, 160, 608, 2432)
    W_E, W_B, W_Bc = bench.work(N, o, v, naux)
    assert W_E == 2.0 * (160 * 161 // 2) * 608 * 608 * 2432
    assert abs(W_E / 1e12 - 23.159) < 1e-3
    assert abs(W_B / 1e12 - 1.973) < 1e-3
    assert W_B - W_Bc == 1.0 * naux * naux * o * v and abs(W_Bc / 1e12 - 1.398) < 1e-3


def test_recorded_traffic_is_read_from_the_profile():
    t = bench.recorded_traffic("rimp2_energy_bf16x3_wide")
    assert isinstance(t, int) and 0.95e9 < t < 400e9   # at least the algorithmic bytes, at most a few re-reads


def test_graft_entry_exposes_build_and_smoke():
    import importlib.util
    spec = importlib.util.spec_from_file_location("graft_entry", os.path.join(ROOT, "__graft_entry__.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    assert callable(mod.build) and callable(mod.smoke)


def test_metric_is_the_headline_metric_and_line_keys_are_unique():
    """The JSON line's "metric" is the headline metric string, and no dict
    literal in bench.py repeats a key (a repeated key silently overwrites)."""
    import ast
    import json
    with open(os.path.join(ROOT, "BASELINE.json")) as f:
        assert bench.METRIC == json.load(f)["metric"]
    with open(os.path.join(ROOT, "bench.py")) as f:
        tree = ast.parse(f.read())
    for node in ast.walk(tree):
        if isinstance(node, ast.Dict):
            keys = [k.value for k in node.keys if isinstance(k, ast.Constant)]
            assert len(keys) == len(set(keys)), keys
