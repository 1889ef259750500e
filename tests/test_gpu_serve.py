"""The request loop: continuous batching with admission by footprint; every
job's result equals the same job run alone, and latencies are reported.

A batch plans small grids onto wider column tiles than a lone instance, so a
job may run a different transform build (fused or staged) in the mix than
alone; the builds order their floating-point operations differently, so the
results agree to rounding: relative L2 per field <= 1e-12 (FP64), 1e-5 (FP32).
"""
import numpy as np
import pytest

import shhd
import serve
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def same_result(job, alone):
    tol = 1e-5 if job.cfg.precision == 32 else 1e-12
    return all(rel_l2(job.result[f], alone[f]) <= tol for f in range(job.cfg.num_fields))


def test_served_stream_results_equal_alone():
    print("[case] served stream, 14 jobs, budget-limited", flush=True)
    jobs = serve.served_stream(14, rate=200.0, steps_range=(3, 40), sizes=(32, 64, 128), seed=3)
    budget = 3 * max(shhd.Batch.footprint(j.cfg) for j in jobs)   # forces queueing
    srv = serve.JobServer(budget_bytes=budget, max_chunk=8)
    try:
        rep = srv.run(jobs)
    finally:
        srv.close()
    p50, p99 = rep.latency_percentiles()
    print(f"    chunks {rep.chunks} peak resident {rep.peak_resident} p50 {p50*1e3:.1f} ms p99 {p99*1e3:.1f} ms",
          flush=True)
    assert rep.peak_resident >= 2
    assert all(j.finished >= j.arrival for j in jobs)
    _, alone = serve.run_sequential(jobs)
    for j, a in zip(sorted(jobs, key=lambda j: (j.arrival, j.job_id)), alone):
        assert same_result(j, a), f"job {j.job_id}: mixed result differs from alone"


@pytest.mark.parametrize("random_start", [False, True])
def test_ensemble_uneven_sizes(random_start):
    print(f"[case] ensemble of 12 uneven quiescent instances random_start={random_start}", flush=True)
    jobs = serve.ensemble(12, nsteps=9, sizes=(32, 64, 128, 256), random_start=random_start)
    srv = serve.JobServer(max_chunk=9)
    try:
        rep = srv.run(jobs)
    finally:
        srv.close()
    assert rep.chunks == 1
    _, alone = serve.run_sequential(jobs)
    for j, a in zip(jobs, alone):
        assert same_result(j, a), f"job {j.job_id}: mixed result differs from alone"
        assert np.linalg.norm(a[1]) > 0, "the noise moves psi"
