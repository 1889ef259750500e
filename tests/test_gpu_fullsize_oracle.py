"""GPU against the FP64 oracle where the served and the 4096^2 Model-H workloads run.

Deterministic runs (kBT = 0; the oracle draws the reference's mt19937_64
stream, the device a Philox stream, so noise is compared in distribution
elsewhere), reference-layout states in and out through the C-ABI:

- Model-H N=2 at 1024^2 (3/2 -> 1536^2), 100 SRK3 steps: FP64 <= 1e-10
  relative L2 per field (the acceptance bar of the path), FP32 form <= 1e-5;
- Model-H N=4 (the any-order row pass built for up to four order parameters,
  which the served stream's N = 4 jobs run) at 256^2 for 20 steps and at
  1024^2 for 3 steps;
- Model-H with N = 3, 5 and 8 distinct order parameters (the builds for up to
  four and up to eight) at 128^2 for 10 steps and N = 8 at 512^2 for 3;
- Model-H N=2 at 2048^2 (3072-point fused builds, the served stream's
  largest grid), 3 steps;
- Model-H N=2 at 1024^2 with forward Euler, 4 steps;
- Model-H N=2 at 4096^2 (6144-point fused builds, the largest single-grid workload),
  one step.

The oracle runs on all host threads it is given; at 1024^2 a step takes about
half a second on 16 threads.  Every case prints its name before touching the
GPU.
"""
import dataclasses

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import rel_l2
from test_gpu_fullsize import check_properties, model_h, synthetic_state

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[case] {name}", flush=True)


def model_h_n(na, n, precision):
    cfg = model_h(na, precision, noise=False)
    return dataclasses.replace(cfg, num_order_parameters=n, mobility=[1.0] * n, mu_k0=[-1.0] * n,
                               mu_k2=[1.0] * n, phi4_u=[1.0] * n)


def oracle_run(cfg64, state, steps):
    o = op.Oracle(cfg64)
    o.set_state(state)
    o.step(steps, 0.0)
    return o.get_state()


def gpu_run(cfg, state, steps):
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(state)
        g.step(steps)
        return g.get_state(), g.info()


def compare(cfg, state, got, ref, tol):
    errs = [rel_l2(got[f], ref[f]) for f in range(cfg.num_fields)]
    print(f"    per-field rel L2 {['%.2e' % e for e in errs]} (tol {tol:.0e})", flush=True)
    for f, e in enumerate(errs):
        assert e <= tol, f"field {f}: rel L2 {e:.3e} > {tol:.1e}"
    check_properties(cfg, state, got, 1e-12 if cfg.precision == 64 else 1e-6)


@pytest.fixture(scope="module")
def model_h_1024_100():
    cfg = model_h(1024, 64, noise=False)
    st = synthetic_state(cfg, seed=12)
    banner("oracle: model-H N=2 1024^2, 100 SRK3 steps")
    return st, oracle_run(cfg, st, 100)


@pytest.mark.parametrize("precision,tol", [(64, 1e-10), (32, 1e-5)])
def test_model_h_1024_100_steps(model_h_1024_100, precision, tol):
    banner(f"model-H N=2 1024^2 100 steps vs oracle fp{precision}")
    st, ref = model_h_1024_100
    cfg = model_h(1024, precision, noise=False)
    got, info = gpu_run(cfg, st, 100)
    assert info.row_fused > 0 and info.forward_column_fused > 0
    compare(cfg, st, got, ref, tol)


@pytest.mark.parametrize("na,steps", [(256, 20), (1024, 3)])
def test_model_h_four_order_parameters(na, steps):
    cfg64 = model_h_n(na, 4, 64)
    st = synthetic_state(cfg64, seed=13)
    banner(f"oracle: model-H N=4 {na}^2, {steps} steps")
    ref = oracle_run(cfg64, st, steps)
    for precision, tol in ((64, 1e-10), (32, 1e-5)):
        banner(f"model-H N=4 {na}^2 {steps} steps vs oracle fp{precision}")
        cfg = model_h_n(na, 4, precision)
        got, info = gpu_run(cfg, st, steps)
        assert info.num_inverse_fields == 7 and info.num_forward_fields == 17
        compare(cfg, st, got, ref, tol)


def model_h_distinct(na, n, precision):
    """N order parameters with a different mobility, quadratic coefficients
    and u for each, so a mix-up of the index alpha anywhere shows."""
    cfg = model_h(na, precision, noise=False)
    return dataclasses.replace(cfg, num_order_parameters=n, mobility=[1.0 + 0.1 * a for a in range(n)],
                               mu_k0=[-1.0 + 0.05 * a for a in range(n)], mu_k2=[1.0 - 0.05 * a for a in range(n)],
                               phi4_u=[1.0 + 0.2 * a for a in range(n)])


@pytest.mark.parametrize("na,n,steps", [(128, 3, 10), (128, 5, 10), (128, 8, 10), (512, 8, 3)])
def test_model_h_up_to_eight_order_parameters(na, n, steps):
    """The row-pass builds for up to four (N = 3: one slot unused) and up to
    eight order parameters (N = 5, 8; the largest the C-ABI accepts), staged
    rows of 192 points and fused rows of 768."""
    cfg64 = model_h_distinct(na, n, 64)
    st = synthetic_state(cfg64, seed=20 + n)
    banner(f"oracle: model-H N={n} {na}^2, {steps} steps")
    ref = oracle_run(cfg64, st, steps)
    for precision, tol in ((64, 1e-10), (32, 1e-5)):
        banner(f"model-H N={n} {na}^2 {steps} steps vs oracle fp{precision}")
        cfg = model_h_distinct(na, n, precision)
        got, info = gpu_run(cfg, st, steps)
        assert info.num_inverse_fields == n + 3 and info.num_forward_fields == 3 * n + 5
        compare(cfg, st, got, ref, tol)


def test_model_h_2048():
    cfg64 = model_h(2048, 64, noise=False)
    st = synthetic_state(cfg64, seed=14)
    banner("oracle: model-H N=2 2048^2, 3 steps")
    ref = oracle_run(cfg64, st, 3)
    for precision, tol in ((64, 1e-10), (32, 1e-5)):
        banner(f"model-H N=2 2048^2 3 steps vs oracle fp{precision}")
        cfg = model_h(2048, precision, noise=False)
        got, info = gpu_run(cfg, st, 3)
        assert info.row_fused > 0 and info.forward_column_fused > 0
        compare(cfg, st, got, ref, tol)


def test_model_h_1024_euler():
    """Forward Euler through the fused 1536-point builds: one stage a step,
    whose update has no base-state term (a graph of three steps, then one
    step outside it)."""
    euler = lambda p: dataclasses.replace(model_h(1024, p, noise=False), integrator=shhd.INTEGRATOR_EULER)  # noqa: E731
    cfg64 = euler(64)
    st = synthetic_state(cfg64, seed=15)
    banner("oracle: model-H N=2 1024^2, 4 Euler steps")
    ref = oracle_run(cfg64, st, 4)
    for precision, tol in ((64, 1e-10), (32, 1e-5)):
        banner(f"model-H N=2 1024^2 4 Euler steps vs oracle fp{precision}")
        cfg = euler(precision)
        got, info = gpu_run(cfg, st, 4)
        assert info.row_fused > 0 and info.forward_column_fused > 0
        compare(cfg, st, got, ref, tol)


def test_model_h_4096_one_step():
    cfg64 = model_h(4096, 64, noise=False)
    st = synthetic_state(cfg64, seed=15)
    banner("oracle: model-H N=2 4096^2, 1 step")
    ref = oracle_run(cfg64, st, 1)
    for precision, tol in ((64, 1e-10), (32, 1e-5)):
        banner(f"model-H N=2 4096^2 1 step vs oracle fp{precision}")
        cfg = model_h(4096, precision, noise=False)
        got, info = gpu_run(cfg, st, 1)
        assert info.row_fused > 0 and info.forward_column_fused > 0
        compare(cfg, st, got, ref, tol)
