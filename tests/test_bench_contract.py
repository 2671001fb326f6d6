"""bench.py's output contract: one JSON line with the metric, roofline and CPU-baseline
fields, and a clear refusal of more than one GPU."""
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode != 0 and "single" in r.stderr and not r.stdout.strip()


@pytest.mark.gpu
def test_bench_json_line():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--steps", "5", "--warmup", "2",
                        "--no-batched", "--no-large", "--no-ddft", "--no-converge", "--no-hipfft"],
                       capture_output=True, text=True, timeout=600, cwd=ROOT)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [l for l in r.stdout.splitlines() if l.startswith("{")]
    assert len(lines) == 1
    d = json.loads(lines[0])
    for key in ("metric", "value", "unit", "n_gpus", "steps", "warmup", "ms_per_step", "higher_is_better", "scaling",
                "vs_baseline", "dtype", "data", "config", "roofline", "cpu_baseline"):
        assert key in d, key
    assert d["n_gpus"] == 1 and d["steps"] == 5 and d["warmup"] == 2 and d["value"] > 0
    assert "workload" in d["config"]
    roof = d["roofline"]
    for key in ("bound", "achieved", "peak", "unit", "frac", "traffic"):
        assert key in roof, key
    assert roof["bound"] == "hbm" and roof["peak"] == 8000.0 and abs(roof["frac"] - roof["achieved"] / 8000.0) < 1e-12
    cpu = d["cpu_baseline"]
    for key in ("value", "unit", "cores", "kind", "sample"):
        assert key in cpu, key
    assert cpu["kind"] in ("port", "reference") and cpu["cores"] == 1
    assert cpu["host"]["nproc"] >= 1 and "model" in cpu["host"]
    assert cpu["all_cores"]["cores"] == cpu["host"]["nproc"] and cpu["all_cores"]["value"] > 0
    assert cpu["gpu_vs_oracle_same_iteration"]["fp64"]["phi_max_rel"] < 1e-10


def test_pmc_summary(tmp_path):
    """--pmc-summary: per kernel and counter, dispatch count, median and maximum over the
    dispatches of the values summed over hardware instances."""
    import sqlite3

    db = tmp_path / "run_results.db"
    con = sqlite3.connect(db)
    con.execute("create table pmc_events (name text, dispatch_id int, counter_name text, counter_value real)")
    rows = [("k_line_pass_reg<float>", 1, "FETCH_SIZE", 10.0), ("k_line_pass_reg<float>", 1, "FETCH_SIZE", 5.0),
            ("k_line_pass_reg<float>", 2, "FETCH_SIZE", 30.0), ("k_line_pass_reg<float>", 3, "FETCH_SIZE", 20.0),
            ("k_other", 4, "FETCH_SIZE", 99.0)]
    con.executemany("insert into pmc_events values (?, ?, ?, ?)", rows)
    con.commit()
    con.close()
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--pmc-summary", str(db), "--pmc-regex",
                        "line_pass"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = r.stdout.strip().splitlines()
    assert len(lines) == 1
    assert "dispatches=3" in lines[0] and "median=20.0" in lines[0] and "max=30.0" in lines[0]
