"""bench.py's host-side contract, checked without a GPU: one GPU only, the
committed rocprofv3 summaries it takes HBM traffic from, and the dpotrf
operation count its FLOP rates are priced with."""
import importlib.util
import subprocess
import sys
from pathlib import Path

REPO = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", REPO / "bench.py")
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(REPO / "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_every_profile_entry_has_its_traffic_line():
    b = _bench()
    for (workload, path, groups) in b.CHAINED_PROFILES:
        traffic, src = b.profiled_traffic(path, True, groups, workload)
        assert traffic is not None and traffic > 0, (workload, path, groups)
        assert (REPO / src).is_file() and src.startswith("profiles/")
    for path in b.PROFILES:
        traffic, src = b.profiled_traffic(path, False, 1, "methanol")
        assert traffic is not None and traffic > 0, path


def test_unprofiled_combinations_report_no_traffic():
    b = _bench()
    assert b.profiled_traffic("lockstep-kc16-32x64", True, 3, "methanol") == (None, None)
    assert b.profiled_traffic("lockstep-kc16", False, 1, "water") == (None, None)


def _cholesky_operations(n: int) -> int:
    """Operations of an unblocked Cholesky, counted step by step: per column
    one square root, a division per entry below the pivot, and a multiply
    and a subtraction per entry of the trailing lower triangle."""
    ops = 0
    for j in range(n):
        below = n - j - 1
        ops += 1 + below + 2 * (below * (below + 1) // 2)
    return ops


def test_potrf_flops_is_the_counted_cholesky():
    b = _bench()
    for n in list(range(1, 33)) + [64, 100, 2000]:
        assert abs(b.potrf_flops(n) - _cholesky_operations(n)) < 1e-9 * max(1, _cholesky_operations(n)), n
