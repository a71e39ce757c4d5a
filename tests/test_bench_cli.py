"""bench.py's command line, without a GPU: one GPU per measurement (several
GPUs are refused with a message before any device is touched), and the
defaults the measurement is quoted on."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_several_gpus_are_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and "--gpus 2" in r.stderr
    assert r.stdout == ""


def test_defaults():
    sys.path.insert(0, ROOT)
    import bench
    argv = sys.argv
    try:
        sys.argv = ["bench.py"]
        args = bench.parse()
    finally:
        sys.argv = argv
    assert args.gpus == 1 and args.steps >= 1 and args.warmup >= 0
    assert args.config == "ct_body"
    assert args.low_energy_correction == 1
    assert args.fp64_stride == 1
