"""Request loop over the batched engine: continuous batching of simulation jobs.

Jobs (model, grid, steps, seed, precision) arrive over time; the loop keeps
every admitted job's state resident in HBM, admits queued jobs in arrival
order whenever the device budget and slots allow (admission by footprint,
`shhd_batch_footprint`), advances the resident jobs in time-bounded chunks of
steps (one launch per kernel, stage and launch shape for every job that
runs, replayed from captured hipGraphs) with the least remaining work first,
and retires each job as soon as its own step count is reached, so jobs that
arrive while others run join at the next chunk boundary.

Each job's result is bit-identical to running it alone (every instance uses
its own Philox counters and the same kernels whatever else is resident).
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np

import shhd


@dataclass
class Job:
    job_id: int
    arrival: float                 # seconds after the start of the stream
    cfg: shhd.SolverConfig
    make_state: Optional[Callable[[], np.ndarray]]   # initial state built on the host at admission, or
    nsteps: int
    ics: Optional[list] = None             # ... initial conditions evaluated on the device (make_state None)
    probes: Optional[list] = None          # (source, index, gx, gy) recorded every step in FP64
    # filled by the server
    slot: int = -1
    remaining: int = 0
    admitted: float = -1.0
    finished: float = -1.0
    result: Optional[np.ndarray] = None
    probe_rows: Optional[np.ndarray] = None
    error: Optional[str] = None            # the job stopped (zero density at step `failed_step`)
    failed_step: int = -1

    @property
    def points(self) -> int:
        return self.cfg.nx_active * self.cfg.ny_active

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


@dataclass
class ServeReport:
    jobs: List[Job]
    makespan: float
    chunks: int
    peak_resident: int
    busy: float = 0.0

    @property
    def point_steps(self) -> float:
        return float(sum(j.points * (j.failed_step if j.error else j.nsteps) for j in self.jobs))

    def throughput(self) -> float:
        """Mpoint*steps/s over the whole stream (first arrival to last finish)."""
        return self.point_steps / self.makespan / 1e6

    def latency_percentiles(self):
        lat = np.array([j.latency for j in self.jobs])
        return float(np.percentile(lat, 50)), float(np.percentile(lat, 99))


class JobServer:
    """Continuous-batching server.

    policy "srpt" (default): every job that has arrived is admitted while the
    device budget and slots allow, and each chunk steps the resident jobs with
    the least remaining work (the smallest, every job within twice its
    remaining work, and further jobs in that order while the chunk's step
    stays under `fill_s`, so cheap jobs share the launches); the others stay
    resident but held.  Chunks are bounded by time (`quantum_s`, from an
    online estimate of the step cost), so new arrivals are ranked at least
    that often.  A short job that arrives behind a long one therefore finishes
    at once, and no job waits behind work larger than its own.
    policy "lockstep": every resident job steps in every chunk (chunk =
    min(max_chunk, smallest remaining step count)).
    """

    def __init__(self, budget_bytes: int = 0, max_chunk: int = 48, use_graphs: bool = True,
                 keep_results: bool = True, clock: Callable[[], float] = time.perf_counter, policy: str = "srpt",
                 quantum_s: float = 0.02, fill_s: float = 2e-4):
        if policy not in ("srpt", "lockstep"):
            raise ValueError(f"unknown policy {policy}")
        self.batch = shhd.Batch(budget_bytes=budget_bytes, use_graphs=use_graphs)
        self.budget = budget_bytes
        self.max_chunk = max_chunk
        self.keep_results = keep_results
        self.clock = clock
        self.policy = policy
        self.quantum = quantum_s
        self.fill = fill_s
        self.sec_per_unit = 6e-10   # step cost per weighted point (calibrated online)

    def close(self):
        self.batch.close()

    @staticmethod
    def _units(job: "Job") -> float:
        """Relative cost of one step of a job: active points x fields, FP32 at
        about 0.6 of FP64 (the measured ratio of the step times)."""
        c = job.cfg
        return c.nx_active * c.ny_active * (c.num_order_parameters + 3) / 5.0 * (1.0 if c.precision == 64 else 0.6)

    def _finish(self, j: "Job", t0: float, done: List["Job"]):
        if self.keep_results:
            j.result = self.batch.get_state(j.slot)
        if j.probes:
            j.probe_rows = self.batch.read_probes(j.slot)
        self.batch.retire(j.slot)
        j.finished = self.clock() - t0
        done.append(j)

    def run(self, jobs: List[Job]) -> ServeReport:
        pending = sorted(jobs, key=lambda j: (j.arrival, j.job_id))
        resident: List[Job] = []
        done: List[Job] = []
        held = set()
        t0 = self.clock()
        chunks = 0
        peak = 0
        busy = 0.0
        qi = 0
        while len(done) < len(jobs):
            now = self.clock() - t0
            # admit arrived jobs in order while they fit
            while qi < len(pending) and pending[qi].arrival <= now:
                j = pending[qi]
                try:
                    j.slot = self.batch.admit(j.cfg, j.make_state() if j.make_state else None)
                except shhd.ShhdError as e:
                    if e.status == shhd.ERR_FULL and resident:
                        break          # wait for memory / slots to free up
                    raise
                for ic in j.ics or []:
                    self.batch.apply_initial_condition(j.slot, ic)
                if j.probes:
                    self.batch.set_probes(j.slot, j.probes, capacity=j.nsteps)
                j.remaining = j.nsteps
                j.admitted = self.clock() - t0
                resident.append(j)
                qi += 1
            peak = max(peak, len(resident))
            if not resident:
                # idle until the next arrival
                wait = pending[qi].arrival - (self.clock() - t0)
                if wait > 0:
                    time.sleep(wait)
                continue
            if self.policy == "lockstep":
                run = resident
                chunk = min(self.max_chunk, min(j.remaining for j in resident))
            else:
                order = sorted(resident, key=lambda j: (j.remaining * self._units(j), j.arrival))
                least = order[0].remaining * self._units(order[0])
                run, step_s = [], 0.0
                for j in order:
                    cost = self.sec_per_unit * self._units(j)
                    if not run or j.remaining * self._units(j) <= 2.0 * least or step_s + cost <= self.fill:
                        run.append(j)
                        step_s += cost
                chunk = min(min(j.remaining for j in run), max(1, int(self.quantum / max(step_s, 1e-9))))
                runs = {j.job_id for j in run}
                for j in resident:   # hold what does not run, release what does
                    if (j.job_id in runs) == (j.slot in held):
                        self.batch.hold(j.slot, j.job_id not in runs)
                        (held.discard if j.job_id in runs else held.add)(j.slot)
            tb = self.clock()
            stopped = {}
            try:
                self.batch.step(chunk)
            except shhd.ShhdError as e:
                if e.status != shhd.ERR_ZERO_DENSITY:
                    raise
                # the instances whose density reached zero stopped at that step;
                # every other job advanced the whole chunk
                stopped = self.batch.failures()
            el = self.clock() - tb
            busy += el
            chunks += 1
            runs = {j.job_id for j in run}
            units = sum(self._units(j) for j in run) * chunk
            if units > 0 and not stopped:
                self.sec_per_unit = 0.7 * self.sec_per_unit + 0.3 * el / units
            still = []
            for j in resident:
                if j.slot in stopped:
                    j.failed_step = stopped[j.slot]
                    j.error = f"compressible velocity encountered zero density. (step {j.failed_step})"
                    held.discard(j.slot)
                    self._finish(j, t0, done)
                    continue
                if j.job_id in runs:
                    j.remaining -= chunk
                if j.remaining == 0:
                    held.discard(j.slot)
                    self._finish(j, t0, done)
                else:
                    still.append(j)
            resident = still
        makespan = max(j.finished for j in jobs) - min(j.arrival for j in jobs)
        return ServeReport(jobs=jobs, makespan=makespan, chunks=chunks, peak_resident=peak, busy=busy)


def run_sequential(jobs: List[Job], clock: Callable[[], float] = time.perf_counter, keep_results: bool = True,
                   durations: Optional[List[float]] = None):
    """The same jobs as plain single-simulation runs, one after another in
    arrival order (final states downloaded only with keep_results, as the
    server does); each job's wall time is appended to `durations` if given."""
    t0 = clock()
    out = []
    for j in sorted(jobs, key=lambda j: (j.arrival, j.job_id)):
        ts = clock()
        with shhd.GpuFluidSolver(j.cfg) as g:
            if j.make_state:
                g.set_state(j.make_state())
            for ic in j.ics or []:
                g.apply_initial_condition(ic)
            g.step(j.nsteps)
            out.append(g.get_state() if keep_results else None)
        if durations is not None:
            durations.append(clock() - ts)
    return clock() - t0, out


def fifo_latencies(jobs: List[Job], durations: List[float]) -> List[float]:
    """Latencies (finish - arrival) of a one-job-at-a-time FIFO server that
    receives the jobs at their arrival times and runs each for its measured
    sequential duration (durations in arrival order, as run_sequential)."""
    t = 0.0
    lat = []
    for j, d in zip(sorted(jobs, key=lambda j: (j.arrival, j.job_id)), durations):
        t = max(t, j.arrival) + d
        lat.append(t - j.arrival)
    return lat


# ---------------------------------------------------------------------------
# synthetic job streams
# ---------------------------------------------------------------------------
def random_state(cfg: shhd.SolverConfig, rng: np.random.Generator, psi_sigma: float = 0.01,
                 kmax: int = 32) -> np.ndarray:
    """rho = 1, psi ~ N(0, psi_sigma^2) and a small momentum, both carried by
    the modes with |kx|, |ky| < kmax (a long-wavelength random field, cheap to
    build on the host for any grid); Hermitian on the kx = 0 column."""
    m2 = cfg.nx * cfg.ny
    st = np.zeros(cfg.state_shape(), np.complex128)
    ny = cfg.ny
    kx_n = min(cfg.nx_active // 2, kmax)
    ky_n = min(cfg.ny_active // 2, kmax)
    nmodes = kx_n * (2 * ky_n - 1)
    for f in range(cfg.num_fields):
        amp = psi_sigma if 1 <= f <= cfg.num_order_parameters else (1e-3 if f > 0 else 0.0)
        if amp == 0.0:
            continue
        scale = amp * m2 / math.sqrt(2.0 * nmodes)
        z = rng.standard_normal((2 * ky_n - 1, kx_n, 2))
        v = (z[..., 0] + 1j * z[..., 1]) * scale
        st[f, :ky_n, :kx_n] = v[:ky_n]
        st[f, ny - ky_n + 1:, :kx_n] = v[ky_n:]
        col = st[f][:, 0]
        col[ny - ky_n + 1:] = np.conj(col[1:ky_n][::-1])
        col[0] = col[0].real
    st[:, 0, 0] = 0
    st[0, 0, 0] = m2
    return st


def model_config(na: int, n: int, precision: int, seed: int, dt: float = 0.005) -> shhd.SolverConfig:
    """Compressible NS (n = 0) or Model-H with n phi4 fields (a=-1, kappa=1, b=1),
    dx = 1, 3/2 padding, SRK3, noise kBT = 1e-3."""
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=dt,
                             num_order_parameters=n, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                             mobility=[1.0] * n, free_energy_kind=shhd.FREE_ENERGY_PHI4 if n else shhd.FREE_ENERGY_NULL,
                             mu_k0=[-1.0] * n, mu_k2=[1.0] * n, phi4_u=[1.0] * n, momentum_advection=True,
                             order_parameter_advection=n > 0, noise_momentum=True, noise_order_parameter=n > 0,
                             kBT=1e-3, seed=seed, precision=precision)


def quiescent_config(na: int, seed: int, precision: int = 64) -> shhd.SolverConfig:
    """Quiescent conserved phi4 field (a=-0.5, kappa=1, b=1, M=1, kBT=0.05),
    SRK3, dt = 0.02 with dx = 1.5 (dx = 1 is beyond the stability bound)."""
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=1.5 * na, ly=1.5 * na, dt=0.02,
                             mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                             free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5], mu_k2=[1.0], phi4_u=[1.0],
                             noise_order_parameter=True, kBT=0.05, seed=seed, precision=precision)


def served_stream(njobs: int, rate: float, steps_range=(10, 200), seed: int = 5,
                  sizes=(256, 512, 1024, 2048), weights=None, fields=(0, 1, 2, 4),
                  precisions=(32, 64)) -> List[Job]:
    """Poisson arrivals at `rate` jobs/s; grid, N, precision drawn per job;
    steps log-uniform in steps_range."""
    rng = np.random.default_rng(seed)
    if weights is None:
        weights = (0.4, 0.3, 0.2, 0.1) if len(sizes) == 4 else (1.0,) * len(sizes)
    t = 0.0
    jobs = []
    for i in range(njobs):
        t += rng.exponential(1.0 / rate)
        na = int(rng.choice(sizes, p=np.array(weights) / np.sum(weights)))
        n = int(rng.choice(fields))
        prec = int(rng.choice(precisions))
        lo, hi = steps_range
        nsteps = int(round(math.exp(rng.uniform(math.log(lo), math.log(hi)))))
        cfg = model_config(na, n, prec, seed=1000 + i)
        sseed = int(rng.integers(1 << 31))
        jobs.append(Job(job_id=i, arrival=t, cfg=cfg, nsteps=nsteps, make_state=None, ics=job_ics(cfg, sseed)))
    return jobs


def job_ics(cfg: shhd.SolverConfig, seed: int):
    """Initial state of a served job, evaluated on the device: uniform density
    1, psi_a from `set order_parameter a equilibrium/gaussian kBT 1e-4 k0 1`
    and momentum from `set momentum all equilibrium/gaussian/compressible
    kBT 1e-4` (per-job seed)."""
    IC = shhd.initial_condition
    ics = [IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=seed, kBT=0.0, value=1.0, pressure_coefficient=1.0)]
    ics += [IC(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, component=a, seed=seed, kBT=1e-4, k0=1.0)
            for a in range(cfg.num_order_parameters)]
    ics += [IC(shhd.IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE, component=d, seed=seed, kBT=1e-4) for d in (0, 1)]
    return ics


def quiescent_ics():
    """Start of an ensemble realization, evaluated on the device: uniform
    density 1 and psi = 0 (`set order_parameter 0 equilibrium/gaussian psi0 0
    kBT 0`); the noise drives it from there."""
    IC = shhd.initial_condition
    return [IC(shhd.IC_EQUILIBRIUM_DENSITY, kBT=0.0, value=1.0, pressure_coefficient=1.0),
            IC(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, component=0, kBT=0.0, value=0.0, k0=1.0)]


def ensemble(ninst: int, nsteps: int, seed0: int = 100, sizes=(64, 128, 256, 512), seed: int = 7,
             low_modes: int = 0, random_start: bool = False) -> List[Job]:
    """Independent quiescent realizations of uneven sizes, all present at t = 0,
    each from psi = 0 set on the device (no host state), or with random_start
    from a long-wavelength random psi built on the host and uploaded at
    admission; with low_modes > 0 each records its lowest Fourier modes of psi
    every step (FP64 probe rows, for the time correlation)."""
    import measures
    rng = np.random.default_rng(seed)
    jobs = []
    for i in range(ninst):
        na = int(rng.choice(sizes))
        cfg = quiescent_config(na, seed=seed0 + i)
        sseed = int(rng.integers(1 << 31))
        probes = None
        if low_modes:
            probes = [(shhd.PROBE_STATE, 1, gx, gy) for gx, gy in measures.lowest_modes(cfg, low_modes)]
        if random_start:
            jobs.append(Job(job_id=i, arrival=0.0, cfg=cfg, nsteps=nsteps, probes=probes,
                            make_state=lambda cfg=cfg, sseed=sseed: random_state(cfg, np.random.default_rng(sseed),
                                                                                 0.1)))
        else:
            jobs.append(Job(job_id=i, arrival=0.0, cfg=cfg, nsteps=nsteps, probes=probes, make_state=None,
                            ics=quiescent_ics()))
    return jobs
