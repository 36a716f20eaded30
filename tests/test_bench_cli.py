"""bench.py's command line on the CPU: it runs on one GPU and says so when
asked for more (no GPU work happens before that check)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_refuses_several_gpus():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""


def test_bench_help_lists_workloads():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--help"], cwd=ROOT,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0
    for w in ("warm", "align", "batch", "cgls", "rotate", "unet"):
        assert w in r.stdout


def test_bench_names_every_forward_design():
    """Every design the autotuner may pick has a kernel name (for the
    roofline's per-kernel timing) and a description in the bench line."""
    sys.path.insert(0, ROOT)
    import bench
    from arec3d_amd.autotune import DESIGNS

    for d in DESIGNS:
        assert d in bench.FORWARD_KERNELS and d in bench.FORWARD_DESIGNS
