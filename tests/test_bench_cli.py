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


def test_bench_kernel_label_from_reported_shape():
    """The kernel named in the bench line is the one the library reports it
    launched (rays per thread, searched stacks, counters, LDS band records)."""
    sys.path.insert(0, str(ROOT))
    import bench
    from rayfactorcl_amd import _lib as L
    t = {"shape_flags": L.RF_SHAPE_STACKS | L.RF_SHAPE_SEARCHED | L.RF_SHAPE_BAND_TILE, "shape_rays_per_thread": 2}
    assert bench.kernel_name(t, True) == "rf::k_queue<2, false, false, 2, true, false, true, false>"
    assert "LDS" in bench.cylinder_form(t)
    t = {"shape_flags": L.RF_SHAPE_STACKS | L.RF_SHAPE_COUNTED, "shape_rays_per_thread": 4}
    assert bench.kernel_name(t, True) == "rf::k_queue<2, false, false, 4, true, true, false, false>"
    assert bench.cylinder_form(t).endswith("a z-range test per band")
    t = {"shape_flags": L.RF_SHAPE_MFMA | L.RF_SHAPE_COUNTED, "shape_rays_per_thread": 0}
    assert bench.kernel_name(t, False) == "rf::k_queue_mfma<false, false>"
    assert bench.cylinder_form(t) == "every test in full"
    t = {"shape_flags": L.RF_SHAPE_STACKS | L.RF_SHAPE_HIERARCHY, "shape_rays_per_thread": 2}
    assert bench.kernel_name(t, True) == "rf::k_queue<2, false, false, 2, true, false, false, true>"
