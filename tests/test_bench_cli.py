"""bench.py's contract, on the host (no GPU): more than one GPU is refused with
a message before anything touches a device, the workloads have the sizes and
physics their descriptions state, and a published baseline exists for none of
them (vs_baseline stays null)."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402
import shhd  # noqa: E402


@pytest.mark.parametrize("gpus", ["2", "8"])
def test_more_than_one_gpu_is_refused(gpus):
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", gpus, "--steps", "1"],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert "one GPU" in r.stderr or "single-GPU" in r.stderr
    assert r.stdout.strip() == ""   # no result line


@pytest.mark.parametrize("name,na,n,padded", [("model_h_1024", 1024, 2, 1536), ("model_h_4096", 4096, 2, 6144),
                                              ("ns_128", 128, 0, 192)])
@pytest.mark.parametrize("precision", [64, 32])
def test_workload_shapes(name, na, n, padded, precision):
    cfg, desc = bench.workload(name, precision)
    assert (cfg.nx_active, cfg.ny_active) == (na, na)
    assert (cfg.nx, cfg.ny) == (padded, padded)
    assert cfg.num_order_parameters == n
    assert cfg.precision == precision
    assert cfg.integrator == shhd.INTEGRATOR_SRK3 and cfg.mode == shhd.MODE_COMPRESSIBLE
    assert cfg.noise_momentum and (n == 0 or cfg.noise_order_parameter)
    assert f"{na}^2" in desc and f"{padded}^2" in desc


def test_unknown_workload_exits():
    with pytest.raises(SystemExit):
        bench.workload("no_such_workload", 64)
