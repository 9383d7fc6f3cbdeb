"""The autotuner's choices persist across processes in a JSON file keyed by device
architecture and geometry (PYRAFT_AMD_TUNE_CACHE).

A first process times the backprojection variants of a plan and of its OSEM subset
range and writes the file; a second process with the same file gets the same variants
and the recorded times without timing again; a file of another format is ignored.
"""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
sys.path.insert(0, %r)
import torch
from pyraft_amd.device import Plan
p = Plan(300, 300, 200)
v, times = p.tune(reps=1)
s = p.radon(torch.rand((300, 300), device="cuda") + 0.5)
p.em(s, 1, 4)
torch.cuda.synchronize()
print("RESULT " + json.dumps({"variant": v, "times": {str(k): t for k, t in times.items()}}))
""" % ROOT


def _child(path):
    env = dict(os.environ, PYRAFT_AMD_TUNE_CACHE=str(path))
    out = subprocess.run([sys.executable, "-c", CHILD], env=env, capture_output=True, text=True, timeout=240)
    assert out.returncode == 0, out.stderr[-2000:]
    line = [l for l in out.stdout.splitlines() if l.startswith("RESULT ")][-1]
    return json.loads(line[len("RESULT "):])


def test_tuning_choices_persist(tmp_path):
    path = tmp_path / "tune.json"
    print("first process", flush=True)
    first = _child(path)
    doc = json.loads(path.read_text())
    keys = list(doc["entries"])
    print("cache keys:", keys, flush=True)
    assert any("|300|300|200|" in k and "|range" not in k for k in keys)
    assert any("|range50" in k for k in keys)  # the OSEM subsets' variant (200 angles / 4)
    entry = [doc["entries"][k] for k in keys if "|range" not in k][0]
    assert entry["variant"] == first["variant"]
    print("second process", flush=True)
    second = _child(path)
    assert second["variant"] == first["variant"]
    # the recorded times (3 decimals), not a new measurement
    for k, t in second["times"].items():
        assert abs(t - entry["times_us"][int(k)]) < 1e-9, (k, t)
    # a file of another format is ignored: the third process times again and rewrites it
    doc["format"] = "something else"
    path.write_text(json.dumps(doc))
    third = _child(path)
    assert json.loads(path.read_text())["format"] != "something else"
    assert 0 <= third["variant"] < 6
