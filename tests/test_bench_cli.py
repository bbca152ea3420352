"""bench.py's host-side behaviour (no GPU): the one-GPU guard and the device-state probe."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""  # no JSON line for a configuration that was not measured


def test_device_state_never_raises():
    import bench

    st = bench.device_state()
    assert st is None or (isinstance(st, dict) and all(isinstance(v, str) for v in st.values()))
