"""The tile-height tuning persists across processes in a JSON file (ORCTM_NH_TUNE_FILE): a first process
times the heights for its size class and writes the choice; the file is valid JSON with one entry per
line, keyed by the device (architecture and CU count), the precision, the size class and the count
class; a second process reads it instead of timing, shown by planting a different height in the file
and finding the solve run with it (every height gives the same bits, so the solution is unchanged).
An empty ORCTM_NH_TUNE_FILE keeps the choices in memory only."""
import json
import os
import subprocess
import sys

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SCRIPT = f"""
import json, sys
sys.path.insert(0, {ROOT!r}); sys.path.insert(0, {os.path.join(ROOT, 'tests')!r})
import numpy as np
from cases import build
from orctm_amd import abi
from orctm_amd.nh import NHSolver, options
g, divg = build("closed")
pnh = np.zeros((g.IE, g.JE, g.KE), order="F")
with NHSolver(g, options(precision=abi.MIXED)) as s:
    st = s.solve(divg, pnh)
print(json.dumps({{"rows": st["tile_rows"], "status": st["status"], "sum": float(np.abs(pnh).sum())}}))
"""


def run(path):
    env = dict(os.environ, ORCTM_NH_TUNE_FILE=str(path))
    r = subprocess.run([sys.executable, "-c", SCRIPT], env=env, capture_output=True, text=True, timeout=180)
    assert r.returncode == 0, r.stdout + r.stderr
    return json.loads(r.stdout.strip().splitlines()[-1])


def test_tuning_is_written_and_read_back(tmp_path):
    print("[tune cache]", flush=True)
    f = tmp_path / "sub" / "tile_tuning.json"
    first = run(f)
    assert first["status"] == "OK" and first["rows"] in (4, 8, 12)
    doc = json.load(open(f))
    entries = doc["entries"]
    mine = [e for e in entries if e["mode"] == 1 and e["count_class"] == 0]
    assert len(mine) == 1 and mine[0]["rows"] == first["rows"] and isinstance(mine[0]["format"], int)
    assert mine[0]["device"].startswith("gfx950/")
    other = 4 if first["rows"] != 4 else 8
    mine[0]["rows"] = other
    with open(f, "w") as out:  # the library's layout: one entry per line
        out.write('{"entries": [\n' + ",\n".join(json.dumps(e) for e in entries) + "\n]}\n")
    second = run(f)
    assert second["status"] == "OK" and second["rows"] == other
    assert second["sum"] == first["sum"]
    empty = tmp_path / "none.json"
    third = run("")
    assert third["status"] == "OK" and not empty.exists()
