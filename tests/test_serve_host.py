"""Host-side serving logic that needs no GPU: the FIFO latency model used as
the sequential reference for the served stream."""
import serve
import shhd


def _job(i, arrival):
    return serve.Job(job_id=i, arrival=arrival, cfg=shhd.SolverConfig(), make_state=None, nsteps=1)


def test_fifo_latencies_closed_form():
    # arrivals 0, 1, 2 s, each run 2 s: finishes at 2, 4, 6 -> waits grow by 1 s per job
    jobs = [_job(0, 0.0), _job(1, 1.0), _job(2, 2.0)]
    assert serve.fifo_latencies(jobs, [2.0, 2.0, 2.0]) == [2.0, 3.0, 4.0]


def test_fifo_latencies_idle_server_and_order():
    # a job arriving after the server went idle starts at its arrival; jobs are
    # taken in arrival order whatever their order in the list
    jobs = [_job(1, 10.0), _job(0, 0.0)]
    assert serve.fifo_latencies(jobs, [1.0, 3.0]) == [1.0, 3.0]
