"""The bench's command-line contract on the host (no device): one GPU only,
and the result line carries the fields the measurement contract names."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused_with_a_message():
    out = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                         capture_output=True, text=True, timeout=60)
    assert out.returncode == 2
    assert "exactly one GPU" in out.stdout + out.stderr


def test_result_line_fields():
    sys.path.insert(0, ROOT)
    import bench
    res = bench._base("m", 0.5, "s", 5, 1, False, {"workload": "w"}, "f64")
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step",
                "higher_is_better", "scaling", "vs_baseline", "dtype", "data", "config"):
        assert key in res
    assert res["n_gpus"] == 1 and res["ms_per_step"] == 500.0 and res["data"] == "synthetic"
    # algorithmic bytes: the FP32 factor halves the factor values, nothing else
    b64 = bench.kernel_bytes(100, 700, 400, False)
    b32 = bench.kernel_bytes(100, 700, 400, True)
    assert b64[0] - b32[0] == 4 * 400 + 4 * 100      # values and the y vector
    assert b64[2] - b32[2] == 4 * 100                # z read by the p update
    assert all(x > 0 for x in b64 + b32)
