"""bench.py's command line on the CPU: the one-GPU contract and the cylinder
test forms it maps onto the library's per_test option."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_bench_cylinder_test_forms():
    sys.path.insert(0, str(ROOT))
    import bench
    assert {k: v[0] for k, v in bench.CYL_FORMS.items()} == {"search": 0, "walk": 2, "full": 1}
