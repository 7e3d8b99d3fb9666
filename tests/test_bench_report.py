"""The bench's report helpers read the committed kernel profile: every
kernel class the roofline can name has a traffic row, and the byte model is
the compulsory-traffic formula (no GPU needed)."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_profile_has_traffic_for_every_kernel_class():
    for cls in bench.CLASS_KERNEL:
        t = bench.measured_traffic(cls)
        assert t is not None and t > 0, cls


def test_unknown_class_has_no_traffic():
    assert bench.measured_traffic("no_such_class") is None


def test_step_bytes_sums_the_per_pass_model():
    """The closed form equals the sum of the per-pass compulsory bytes it is
    built from (keys, sort, reorder, cell table, external forces, K x
    (predict + density + force), integration)."""
    n, nf, c = 1000, 800, 64
    for K in (3, 12):
        passes = [20 * n, 12 * n, 76 * n, 4 * n + 8 * c, 40 * n + 36 * nf, 96 * nf]
        per_iter = [80 * nf, 16 * n + 12 * nf, 20 * n + 16 * nf]
        assert bench.step_bytes(n, nf, c, K) == sum(passes) + K * sum(per_iter)


def test_more_than_one_gpu_is_refused_with_a_message():
    """The workload is one scene on one GPU: a request for several GPUs ends
    with exit status 2 and says why, before any device is touched."""
    import subprocess
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""
