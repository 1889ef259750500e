#!/usr/bin/env python3
"""Throughput of the MI355X SRK3 step path, in the reference's metric.

Default workload (the headline configuration): Model-H-like compressible flow
with N = 2 conserved phi4 scalars on a 1024 x 1024 active grid (dx = 1, 3/2
padding to 1536^2), phi4 a = -1, kappa = 1, b = 1, M = 1, eta = zeta = 0.5,
cT^2 = 1, rho0 = 1, advection on, SRK3.  The stated dt = 0.01 is beyond the
explicit SRK3 stability bound at dx = 1, so dt = 0.005 is used.  Noise is on
with kBT = 1e-3 (in-kernel Philox), so every stage draws its stochastic terms.

`value` is active-grid Mpoint * SRK3-steps / s on one GPU with the state
resident in HBM; `roofline` is the dominant kernel's algorithmic HBM bytes per
launch over its mean launch time (HIP events on the solver's stream);
`cpu_baseline` times the FP64 CPU oracle on a bounded sample of the same
workload on the host cores.  `model_h_4096` measures the same model at 4096^2
(3/2 -> 6144^2), FP64 and the FP32 form, each with its own roofline.
"""
import argparse
import json
import math
import os
import shutil
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.join(ROOT, "spectral-hohenberg-halperin-dynamics-2d_amd")
sys.path.insert(0, PKG)

import shhd  # noqa: E402

HBM_PEAK_GBS = 8000.0


def workload(name, precision):
    if name == "model_h_1024":
        na, n, dt = 1024, 2, 0.005
        desc = "model-H compressible SRK3, N=2 phi4, 1024^2 active (3/2 -> 1536^2), dx=1, dt=0.005, kBT=1e-3"
    elif name == "model_h_4096":
        na, n, dt = 4096, 2, 0.005
        desc = "model-H compressible SRK3, N=2 phi4, 4096^2 active (3/2 -> 6144^2), dx=1, dt=0.005, kBT=1e-3"
    elif name == "ns_128":
        na, n, dt = 128, 0, 0.05
        desc = "compressible NS SRK3, N=0, 128^2 active (3/2 -> 192^2), dx=1, dt=0.05, kBT=1e-4"
    elif name == "incompressible_4096":
        cfg = shhd.SolverConfig(nx_active=4096, ny_active=4096, dealias="3/2", dt=1e-4,
                                mode=shhd.MODE_INCOMPRESSIBLE, eta=1e-4, momentum_advection=True,
                                noise_momentum=True, kBT=1e-8, seed=3, precision=precision)
        return cfg, "incompressible SRK3, 4096^2 active (3/2 -> 6144^2), L=2pi, eta=1e-4, dt=1e-4, kBT=1e-8"
    else:
        raise SystemExit(f"unknown workload {name}")
    if n == 0:
        cfg = shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=dt,
                                pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True,
                                noise_momentum=True, kBT=1e-4, seed=1, precision=precision)
    else:
        cfg = shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=dt,
                                num_order_parameters=n, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                                mobility=[1.0] * n, free_energy_kind=shhd.FREE_ENERGY_PHI4,
                                mu_k0=[-1.0] * n, mu_k2=[1.0] * n, phi4_u=[1.0] * n,
                                momentum_advection=True, order_parameter_advection=True,
                                noise_momentum=True, noise_order_parameter=True, kBT=1e-3, seed=2,
                                precision=precision)
    return cfg, desc


def initial_state(cfg, seed=2):
    """rho = 1 + small fluctuations, psi ~ 0.01 Gaussian noise, small momentum;
    Hermitian and masked (as every reference initial condition)."""
    rng = np.random.default_rng(seed)
    mask = shhd.active_mask(cfg)
    m2 = cfg.nx * cfg.ny
    st = np.zeros(cfg.state_shape(), np.complex128)
    n_act = mask.sum()
    for f in range(cfg.num_fields):
        amp = 0.01 if 1 <= f <= cfg.num_order_parameters else 1e-3
        z = (rng.standard_normal(mask.shape) + 1j * rng.standard_normal(mask.shape)) * mask
        # physical rms ~ amp: |psi_k| ~ amp * Nc / sqrt(2 * active modes)
        st[f] = z * (amp * m2 / math.sqrt(2.0 * n_act))
    st[:, 0, 0] = 0
    st[0, 0, 0] = m2
    if cfg.mode == shhd.MODE_INCOMPRESSIBLE:
        kx, ky, _ = shhd.wavenumbers(cfg)
        k2 = kx ** 2 + ky ** 2
        n = cfg.num_order_parameters
        with np.errstate(invalid="ignore", divide="ignore"):
            tr = np.where(k2 > 0, (-ky * st[n + 1] + kx * st[n + 2]) / k2, 0)
        st[n + 1] = -ky * tr
        st[n + 2] = kx * tr
    return shhd.hermitian_project(st, cfg)


def device_initial_conditions(name):
    """The `set` commands of the configurations that name reference styles,
    evaluated on the device: 128^2 compressible Navier-Stokes, equilibrium/gaussian
    density and momentum (seed 1, kBT 1e-4, rho0 1, cT 1); 4096^2 incompressible, the
    k^4 exp(-2 (k/8)^2) spectrum,
    urms 1, seed 3, rho0 1.  None: the synthetic state of initial_state()."""
    IC = shhd.initial_condition
    if name == "ns_128":
        return [IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=1, kBT=1e-4, value=1.0, pressure_coefficient=1.0)] + \
               [IC(shhd.IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE, component=d, seed=1, kBT=1e-4) for d in (0, 1)]
    if name == "incompressible_4096":
        return [IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=3, kBT=0.0, value=1.0, pressure_coefficient=1.0)] + \
               [IC(shhd.IC_SPECTRUM_K4, component=d, seed=3, n0=8.0, urms=1.0) for d in (0, 1)]
    if name == "model_h_4096":
        # equilibrium fluctuations set on the device (a 4096^2 host state is
        # slow to build and upload): rho = 1, psi and j Gaussian, seed 2
        return ([IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=2, kBT=1e-3, value=1.0, pressure_coefficient=1.0)]
                + [IC(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, component=a, seed=2, kBT=1e-4, k0=1.0) for a in (0, 1)]
                + [IC(shhd.IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE, component=d, seed=2, kBT=1e-3) for d in (0, 1)])
    return None


def fp32_error(cfg64, ics, steps=20):
    """Relative L2 error of the FP32 form against FP64 after `steps`
    deterministic steps (noise off: the two forms draw different Gaussians)
    from the same initial state, over all fields."""
    import dataclasses
    out = {}
    for prec in (64, 32):
        c = dataclasses.replace(cfg64, precision=prec, noise_momentum=False, noise_order_parameter=False)
        with shhd.GpuFluidSolver(c) as g:
            if ics:
                for ic in ics:
                    g.apply_initial_condition(ic)
            else:
                g.set_state(initial_state(c))
            g.step(steps)
            out[prec] = g.get_state()
    d = out[32] - out[64]
    per_field = [float(np.linalg.norm(d[f]) / max(np.linalg.norm(out[64][f]), 1e-300)) for f in range(d.shape[0])]
    return {"steps": steps, "rel_l2": float(np.linalg.norm(d) / np.linalg.norm(out[64])), "per_field": per_field}


def time_gpu(cfg, steps, warmup, ics=None):
    """Wall time of `steps` steps (state resident, nothing else on the stream),
    then the per-kernel launch times from HIP events around each launch over
    a second run of the same steps (kept out of the first so the events cost
    the measured rate nothing)."""
    with shhd.GpuFluidSolver(cfg) as g:
        if ics:
            for ic in ics:
                g.apply_initial_condition(ic)
        else:
            g.set_state(initial_state(cfg))
        g.step(warmup, 0.0)
        t0 = time.perf_counter()
        g.step(steps, warmup * cfg.dt)        # returns after the device finished
        t1 = time.perf_counter()
        g.set_kernel_timing(True)
        g.step(steps, (warmup + steps) * cfg.dt)
        kt = g.kernel_timing()
        kb = g.kernel_bytes()
        info = g.info()
        out = g.get_state()
    assert np.all(np.isfinite(out)), "non-finite state"
    return t1 - t0, kt, kb, info


def measured(cfg, steps, warmup, ics, workload_key, precision):
    """Throughput and the dominant kernel's roofline of one workload."""
    el, kt, kb, info = time_gpu(cfg, steps, warmup, ics)
    pts = cfg.nx_active * cfg.ny_active
    dom = max(kt, key=lambda k: kt[k][0])
    ms, n = kt[dom]
    achieved = kb[dom] / (ms / 1e3 / max(n, 1)) / 1e9
    kernels = {k: {"avg_ms": kt[k][0] / max(kt[k][1], 1), "launches": kt[k][1],
                   "bytes_per_launch": kb[k], "GBs": kb[k] / (kt[k][0] / 1e3 / max(kt[k][1], 1)) / 1e9
                   if kt[k][1] else None} for k in kt}
    traffic = None
    prof = os.path.join(ROOT, "profiles", "hbm_traffic.json")
    if os.path.exists(prof):
        try:
            traffic = json.load(open(prof)).get(workload_key, {}).get(f"fp{precision}", {}).get(dom)
        except Exception:
            traffic = None
    return {"value": pts * steps / el / 1e6, "ms_per_step": el * 1e3 / steps, "info": info,
            "roofline": {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                         "frac": achieved / HBM_PEAK_GBS, "traffic": traffic},
            "kernels": kernels,
            "step_GBs": sum(kb[k] * kt[k][1] for k in kt) / max(steps, 1) / (el / steps) / 1e9}


def note(msg):
    """A progress line on stderr (the result is the one JSON line on stdout)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def model_h_4096(steps, warmup, tune):
    """Model-H N=2 at 4096^2 (3/2 -> 6144^2), FP64 and the FP32 form: the
    configuration the path's throughput target is quoted on."""
    out = {}
    for prec in (64, 32):
        cfg, desc = workload("model_h_4096", prec)
        note(f"4096^2 fp{prec}")
        shapes = shhd.autotune(cfg, steps=2) if tune else None
        ics = device_initial_conditions("model_h_4096")
        m = measured(cfg, steps, warmup, ics, "model_h_4096", prec)
        m.pop("info")
        out[f"fp{prec}"] = {"workload": desc, "dtype": f"f{prec}", "launch_shapes": shapes, **m}
    out["fp32"]["speedup_vs_fp64"] = out["fp32"]["value"] / out["fp64"]["value"]
    out["fp32"]["error_vs_fp64"] = fp32_error(workload("model_h_4096", 64)[0],
                                              device_initial_conditions("model_h_4096"), steps=10)
    return out


def ensemble_bench(ninst):
    """Batched ensemble (the quiescent-ensemble workload): ninst quiescent phi4 realizations
    with sizes drawn from {64,128,256,512}^2, each from psi = 0 set on the
    device, 100 SRK3 steps each (the stated 10^4 scaled down), dx = 1.5 for
    stability; all advanced together (one launch per kernel, stage and launch
    shape, hipGraph replay) against the same runs done one after another with
    single handles."""
    import serve
    import measures
    jobs = serve.ensemble(ninst, nsteps=100, low_modes=16)
    srv = serve.JobServer(max_chunk=100, keep_results=False, policy="lockstep")
    try:
        rep = srv.run(jobs)
    finally:
        srv.close()
    # time correlation of the 16 lowest modes of every realization, from its FP64 probe rows
    c = np.array([measures.time_correlation(j.probe_rows, 10) for j in jobs])   # (inst, lag, mode)
    decay = float(np.mean(c[:, 10].real / c[:, 0].real))
    for j in jobs:
        j.probes = None
    seq_s, _ = serve.run_sequential(jobs, keep_results=False)   # neither side downloads results
    pts = rep.point_steps
    return {"instances": ninst, "steps": 100, "Mpoint_steps_s": pts / rep.makespan / 1e6,
            "sequential_Mpoint_steps_s": pts / seq_s / 1e6, "speedup_vs_sequential": seq_s / rep.makespan,
            "batched_s": rep.makespan, "sequential_s": seq_s,
            "low_mode_correlation": {"modes": 16, "rows_per_instance": int(jobs[0].probe_rows.shape[0]),
                                     "mean_C10_over_C0": decay}}


def served_bench(njobs):
    """Served stream (the mixed served-job workload, scaled down): njobs with Poisson
    arrivals at 20 jobs/s, grids {256,512,1024,2048}^2 (weights .4/.3/.2/.1),
    N in {0,1,2,4}, FP32 or FP64, steps log-uniform in [10, 200] (the stated
    [1000, 20000] / 100); continuous batching with admission by footprint,
    least remaining work first in 20 ms chunks (the others held resident)."""
    import serve
    jobs = serve.served_stream(njobs, rate=20.0, steps_range=(10, 200))
    srv = serve.JobServer(keep_results=False, policy="srpt")
    try:
        rep = srv.run(jobs)
    finally:
        srv.close()
    p50, p99 = rep.latency_percentiles()
    durations = []
    seq_s, _ = serve.run_sequential(jobs, keep_results=False, durations=durations)   # neither side downloads
    fifo = np.array(serve.fifo_latencies(jobs, durations))
    pts = rep.point_steps
    return {"jobs": njobs, "arrival_rate_per_s": 20.0, "policy": "srpt", "p50_s": p50, "p99_s": p99,
            "Mpoint_steps_s": pts / rep.makespan / 1e6, "busy_Mpoint_steps_s": pts / rep.busy / 1e6,
            "sequential_Mpoint_steps_s": pts / seq_s / 1e6, "speedup_vs_sequential": seq_s / rep.busy,
            "sequential_fifo_p50_s": float(np.percentile(fifo, 50)),
            "sequential_fifo_p99_s": float(np.percentile(fifo, 99)),
            "makespan_s": rep.makespan, "chunks": rep.chunks, "peak_resident": rep.peak_resident}


def reference_toolchain_probe():
    """What the reference binary would need on this host (src/fourier_transform.h
    includes heffte.h; the build links FFTW and MPI): mpirun, fftw3.h, heFFTe."""
    def header(name):
        for d in ("/usr/include", "/usr/local/include", "/opt/conda/include", "/opt/rocm/include"):
            if os.path.exists(os.path.join(d, name)):
                return os.path.join(d, name)
        return None
    found = {"mpirun": shutil.which("mpirun"), "fftw3.h": header("fftw3.h"), "heffte.h": header("heffte.h")}
    found["reference_buildable"] = all(found[k] for k in ("mpirun", "fftw3.h", "heffte.h"))
    return found


def host_cpus():
    """Logical CPUs of the host (nproc), the ones this process may run on, and
    the threads the oracle uses: OMP_NUM_THREADS when set (the host's CPU share
    for this job), else every CPU the process may run on."""
    total = os.cpu_count() or 1
    try:
        allowed = len(os.sched_getaffinity(0))
    except Exception:
        allowed = total
    env = os.environ.get("OMP_NUM_THREADS", "")
    threads = int(env) if env.isdigit() and int(env) > 0 else allowed
    return total, allowed, min(threads, allowed)


def cpu_baseline(cfg_gpu, budget_s):
    """FP64 CPU oracle (a restatement of the reference algorithm, own FFT,
    OpenMP over rows/columns) on the same workload: 1 untimed step, then whole
    steps until the time budget is spent."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import oracle_py as op
    nproc, allowed, threads = host_cpus()
    used = op.set_threads(threads)
    cfg = shhd.SolverConfig(**{**cfg_gpu.__dict__, "precision": 64})
    o = op.Oracle(cfg)
    o.set_state(initial_state(cfg))
    o.step(1, 0.0)
    n = 0
    t0 = time.perf_counter()
    while True:
        o.step(1, (n + 1) * cfg.dt)
        n += 1
        el = time.perf_counter() - t0
        if el >= budget_s or n >= 50:
            break
    pts = cfg.nx_active * cfg.ny_active
    return {"value": pts * n / el / 1e6, "unit": "Mpoint*SRK3-steps/s", "cores": used, "kind": "port",
            "sample": f"{n} SRK3 steps of the same workload after 1 untimed step, {el:.1f} s, "
                      f"FP64 oracle restatement, {used} OpenMP threads",
            "host": {"nproc": nproc, "cpus_allowed": allowed, "omp_num_threads": os.environ.get("OMP_NUM_THREADS")},
            "reference_toolchain": reference_toolchain_probe()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="model_h_1024")
    ap.add_argument("--precision", type=int, default=64, choices=(32, 64))
    ap.add_argument("--no-fp32", action="store_true", help="skip the FP32-form side measurement")
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="default launch shapes instead of the autotuned ones")
    ap.add_argument("--launch-shapes", default="", metavar="CI,CF,NT[,CT,RT,PB[,TB[,FW[,RW]]]]",
                    help="use these launch shapes (inverse tile, forward tile, row lanes, column and row "
                         "transform choice, log2 row-to-column plane block, column-to-row plane block, "
                         "forward-column and row wave groups, as in shhd_tuning) instead of tuning")
    ap.add_argument("--no-serve", action="store_true", help="skip the ensemble and served-stream measurements")
    ap.add_argument("--serve-jobs", type=int, default=200)
    ap.add_argument("--ensemble", type=int, default=256)
    ap.add_argument("--no-4096", action="store_true", help="skip the 4096^2 FP64/FP32 measurement")
    ap.add_argument("--steps-4096", type=int, default=60)
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU: every configuration of this path fits one MI355X and runs are "
              "single-GPU only; multi-GPU runs would shard by instance.", file=sys.stderr)
        sys.exit(2)

    cfg, desc = workload(args.workload, args.precision)
    # launch shapes for this problem shape, measured before the timed region
    if args.launch_shapes:
        vals = [int(v) for v in args.launch_shapes.split(",")] + [0, 0, 0, 0, 0, 0]
        c1, c3, rnt, ct, rt, pb, tb, fw, rw = vals[:9]
        shhd.set_launch_shapes(cfg, c1, c3, rnt, ct, rt, pb, tb, fw, rw)
        tuning = {"inverse_column_tile": c1, "forward_column_tile": c3, "row_lanes": rnt, "column_transform": ct,
                  "row_transform": rt, "plane_block": pb, "inverse_plane_block": tb, "forward_column_waves": fw,
                  "row_waves": rw, "given": True}
    else:
        note(f"{args.workload} fp{args.precision}: launch shapes")
        tuning = shhd.autotune(cfg, steps=4) if not args.no_tune else None
    note(f"{args.workload} fp{args.precision}: {args.steps} steps")
    ics = device_initial_conditions(args.workload)
    m = measured(cfg, args.steps, args.warmup, ics, args.workload, args.precision)
    info = m["info"]
    value = m["value"]
    pts = cfg.nx_active * cfg.ny_active

    result = {
        "metric": "Mpoint*SRK3-steps/s",
        "value": value,
        "unit": "Mpoint*SRK3-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == 64 else "f32",
        "data": ("device initial conditions of the configuration's `set` commands (seeded)" if ics else
                 "synthetic (rho=1+1e-3 noise, psi ~ 1e-2 Gaussian, small momentum; seeded)"),
        "config": {"workload": desc, "grid": [cfg.nx_active, cfg.ny_active], "padded": [cfg.nx, cfg.ny],
                   "order_parameters": cfg.num_order_parameters, "integrator": "srk3",
                   "F_in": info.num_inverse_fields, "F_out": info.num_forward_fields,
                   "F_s": info.num_updated_fields, "parallelism": "single GPU",
                   "launch_shapes": tuning},
        "roofline": m["roofline"],
        "kernels": m["kernels"],
        "step_GBs": m["step_GBs"],
    }
    if args.precision == 64 and not args.no_fp32:
        cfg32, _ = workload(args.workload, 32)
        note(f"{args.workload} fp32")
        tuning32 = shhd.autotune(cfg32, steps=4) if not args.no_tune else None
        m32 = measured(cfg32, args.steps, args.warmup, ics, args.workload, 32)
        result["fp32"] = {"value": m32["value"], "ms_per_step": m32["ms_per_step"],
                          "speedup_vs_fp64": m32["value"] / value, "launch_shapes": tuning32,
                          "roofline": m32["roofline"], "kernels": m32["kernels"],
                          "error_vs_fp64": fp32_error(cfg, ics)}
    if not args.no_4096 and args.workload == "model_h_1024" and args.precision == 64:
        result["model_h_4096"] = model_h_4096(args.steps_4096, max(2, args.warmup // 4), not args.no_tune)
    if not args.no_serve:
        note("ensemble")
        result["ensemble"] = ensemble_bench(args.ensemble)
        note("served stream")
        result["served"] = served_bench(args.serve_jobs)
        result["p50_latency_s"] = result["served"]["p50_s"]
        result["p99_latency_s"] = result["served"]["p99_s"]
    if not args.no_cpu:
        note("cpu baseline")
        result["cpu_baseline"] = cpu_baseline(cfg, args.cpu_budget)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()
