"""Zero density stops a run where the reference's throw would leave it.

The reference raises "compressible velocity encountered zero density." inside
the failing right-hand side (src/fcalculator_workspace.cc:126-128), before any
stage writes the state (time_integrator_srk3_compressible.h:99-147 writes `u`
only at the third stage), so the state stays at the failing step's start.  A
multi-step call must therefore stop at that step with the state unchanged
since, and a batch must stop only the failing instance.
"""
import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import random_active_state, rel_l2

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[case] {name}", flush=True)


def ns_cfg(na=16, precision=64):
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=0.01,
                             pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True,
                             precision=precision)


def zero_density_state(cfg, rng):
    st = random_active_state(cfg, rng, amp={1: 0.01, 2: 0.01})
    st[0] = 0.0   # uniform density value 0
    return st


@pytest.mark.parametrize("precision", [64, 32])
def test_multi_step_call_stops_at_failing_step(precision):
    banner(f"zero density: 5-step call stops at step 0, fp{precision}")
    cfg = ns_cfg(precision=precision)
    rng = np.random.default_rng(3)
    bad = zero_density_state(cfg, rng)
    with pytest.raises(op.OracleError, match="zero density"):   # the oracle throws at the first step
        o = op.Oracle(ns_cfg())
        o.set_state(bad)
        o.step(5)
    good = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.01, 2: 0.01})
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(bad)
        with pytest.raises(shhd.ShhdError, match="compressible velocity encountered zero density") as ei:
            g.step(5)
        assert ei.value.status == shhd.ERR_ZERO_DENSITY
        after = g.get_state()
        assert np.all(np.isfinite(after))
        assert np.array_equal(after, shhd.hermitian_project(bad, cfg)) or rel_l2(after, bad) <= 1e-7
        # the handle stays usable: a valid state steps normally afterwards
        g.set_state(good)
        g.step(5)
        got = g.get_state()
    o = op.Oracle(ns_cfg())
    o.set_state(good)
    o.step(5)
    ref = o.get_state()
    assert rel_l2(got, ref) <= (1e-10 if precision == 64 else 1e-5)


def test_batch_stops_only_the_failing_instance():
    banner("zero density: batch stops one instance, the others advance")
    rng = np.random.default_rng(4)
    cfg_a = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=32.0, ly=32.0, dt=0.005,
                              num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                              mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-1.0, -1.0],
                              mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0], momentum_advection=True,
                              order_parameter_advection=True, noise_momentum=True, noise_order_parameter=True,
                              kBT=1e-3, seed=9)
    st_a = random_active_state(cfg_a, rng, amp={0: 0.01, 1: 0.2, 2: 0.2, 3: 0.01, 4: 0.01})
    cfg_b = ns_cfg(na=16)
    st_b = zero_density_state(cfg_b, rng)

    with shhd.Batch() as alone:
        sa = alone.admit(cfg_a, st_a)
        alone.step(12)
        ref_a = alone.get_state(sa)

    with shhd.Batch() as b:
        sa = b.admit(cfg_a, st_a)
        b.step(4)
        sb = b.admit(cfg_b, st_b)
        with pytest.raises(shhd.ShhdError, match="zero density") as ei:
            b.step(5)
        assert ei.value.status == shhd.ERR_ZERO_DENSITY
        assert f"slot {sb}" in str(ei.value)
        assert b.failures() == {sb: 0}
        b.step(3)                     # no new failure: the stopped instance stays stopped
        assert b.failures() == {sb: 0}
        got_a = b.get_state(sa)
        got_b = b.get_state(sb)
        b.retire(sb)
        assert b.failures() == {}
        b.step(1)
    assert np.array_equal(got_a, ref_a), "healthy instance must compute what it computes without the failing one"
    assert np.array_equal(got_b, shhd.hermitian_project(st_b, cfg_b))
