"""bench.py's output contract: one JSON line with the metric, the whole-job value, the timing
fields, the workload config, the roofline of the dominant kernel and the CPU baseline; and the
refusal of a multi-GPU request (no GPU needed)."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BENCH = os.path.join(ROOT, "bench.py")


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, BENCH, "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""


@pytest.mark.gpu
@pytest.mark.timeout(600)
def test_bench_line_contract():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    print("case bench.py --steps 5 --warmup 2 --no-extras --cpu-seconds 0.5", flush=True)
    r = subprocess.run([sys.executable, BENCH, "--steps", "5", "--warmup", "2", "--no-extras", "--cpu-seconds", "0.5"],
                       capture_output=True, text=True, timeout=540)
    assert r.returncode == 0, r.stderr[-3000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.strip()]
    assert len(lines) == 1, r.stdout[-2000:]
    d = json.loads(lines[0])
    for k in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
              "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert k in d, k
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 2
    assert d["unit"] == "Gpix/s" and d["higher_is_better"] is True and d["dtype"] == "f16"
    assert d["config"]["workload"].startswith("c3: 3840x2160")
    # value is pixels per second of the timed steps: one 4K frame per step
    assert d["value"] > 0
    assert abs(d["value"] - 3840 * 2160 / (d["ms_per_step"] * 1e-3) / 1e9) <= 0.01 * d["value"]
    rf = d["roofline"]
    assert rf["bound"] == "hbm" and rf["unit"] == "GB/s" and rf["peak"] == 8000.0
    assert abs(rf["frac"] - rf["achieved"] / rf["peak"]) < 0.01
    assert abs(rf["achieved"] - 44 * 3840 * 2160 / (d["ms_per_step"] * 1e-3) / 1e9) <= 0.01 * rf["achieved"]
    cb = d["cpu_baseline"]
    assert cb["value"] > 0 and cb["cores"] >= 1 and cb["kind"] in ("port", "reference") and cb["sample"]
    print("bench line ok:", d["value"], "Gpix/s", flush=True)
