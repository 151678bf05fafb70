"""bench.py's contract parts that need no GPU: a run on several GPUs is refused with a
message, and the HBM traffic figures bench.py reports come from the committed counter
summaries, consistent with their counters and calibration."""
from __future__ import annotations

import importlib.util
import json
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _bench():
    spec = importlib.util.spec_from_file_location("bench_contract", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_several_gpus_are_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no JSON line for a run that did not happen


@pytest.mark.parametrize("name,n", [("k_nve_fused_n262144_traffic.json", 262144),
                                    ("k_bfire_forces_rigid_n65536_traffic.json", 65536)])
def test_traffic_figures_follow_their_counters(name, n):
    b = _bench()
    with open(os.path.join(ROOT, "profiles", name)) as f:
        d = json.load(f)
    got = b.traffic_from_profiles(n, name)
    # KiB per dispatch, each counter scaled by what it reads for a copy of known size
    want = (d["fetch_kib"] * d["fetch_scale"] + d["write_kib"] * d["write_scale"]) * 1024.0
    assert got == pytest.approx(want, rel=1e-6)
    assert os.path.exists(os.path.join(ROOT, d["source"]))
    # a summary for another size is not used
    assert b.traffic_from_profiles(n + 64, name) is None
