"""Batched instances (shhd_batch_*): every instance computes exactly what it
computes alone, whatever else is resident, with or without graph replay."""
import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import rel_l2, random_active_state

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[case] {name}", flush=True)


def model_h(na, precision=64, seed=2, noise=True, n=2):
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=float(na), ly=float(na), dt=0.005,
                             num_order_parameters=n, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                             mobility=[1.0] * n, free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-1.0] * n,
                             mu_k2=[1.0] * n, phi4_u=[1.0] * n, momentum_advection=True,
                             order_parameter_advection=n > 0, noise_momentum=noise, noise_order_parameter=noise and n > 0,
                             kBT=1e-3, seed=seed, precision=precision)


def quiescent(na, seed, precision=64):
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=1.5 * na, ly=1.5 * na, dt=0.02,
                             mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                             free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-0.5], mu_k2=[1.0], phi4_u=[1.0],
                             noise_order_parameter=True, kBT=0.05, seed=seed, precision=precision)


def state_for(cfg, seed):
    rng = np.random.default_rng(seed)
    amp = {0: 0.01}
    for f in range(1, cfg.num_order_parameters + 1):
        amp[f] = 0.2
    if cfg.mode != shhd.MODE_QUIESCENT:
        amp[cfg.num_order_parameters + 1] = 0.01
        amp[cfg.num_order_parameters + 2] = 0.01
    return random_active_state(cfg, rng, amp=amp)


def alone(cfg, st, nsteps, graphs):
    with shhd.Batch(use_graphs=graphs) as b:
        s = b.admit(cfg, st)
        b.step(nsteps)
        return b.get_state(s)


@pytest.mark.parametrize("graphs", [False, True])
def test_alone_equals_mix(graphs):
    banner(f"alone vs mix graphs={graphs}")
    cfgs = [quiescent(64, 100), quiescent(128, 101), quiescent(32, 102, precision=32), model_h(32, seed=5),
            model_h(64, precision=32, seed=6), quiescent(256, 103)]
    states = [state_for(c, i) for i, c in enumerate(cfgs)]
    with shhd.Batch(use_graphs=graphs) as b:
        slots = [b.admit(c, s) for c, s in zip(cfgs, states)]
        b.step(7)
        mixed = [b.get_state(s) for s in slots]
    for c, s, m in zip(cfgs, states, mixed):
        a = alone(c, s, 7, graphs)
        assert np.array_equal(a, m), f"{c.nx_active} fp{c.precision}: alone != mix"


def test_batch_equals_single_handle_and_oracle():
    banner("batch vs single handle vs oracle (kBT=0)")
    cfg = model_h(32, noise=False)
    st = state_for(cfg, 9)
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.step(20)
        single = g.get_state()
    batched = alone(cfg, st, 20, True)
    assert np.array_equal(single, batched)
    o = op.Oracle(cfg)
    o.set_state(st)
    o.step(20)
    assert rel_l2(batched, o.get_state()) < 1e-10


def test_graph_replays_match_eager():
    """Several graph replays, with eager work and host synchronisation between
    them, each checked against the eager run."""
    banner("graph replays vs eager")
    cfgs = [quiescent(64, 7), model_h(32, seed=8)]
    states = [state_for(c, 20 + i) for i, c in enumerate(cfgs)]
    with shhd.Batch(use_graphs=True) as bg, shhd.Batch(use_graphs=False) as be:
        sg = [bg.admit(c, s) for c, s in zip(cfgs, states)]
        se = [be.admit(c, s) for c, s in zip(cfgs, states)]
        for chunk in (3, 6, 4, 3, 5):
            bg.step(chunk)
            be.step(chunk)
            for a, b in zip(sg, se):
                assert np.array_equal(bg.get_state(a), be.get_state(b))


@pytest.mark.parametrize("graphs", [True, False])
def test_step_splits_equal_one_call(graphs):
    """Steps taken in calls of any length (single steps, as a host driving
    the solver once per step does, and runs that do not start at the buffer
    rotation the launch tables were built for) give bit for bit the state and
    step flux of one call over all the steps."""
    banner(f"step splits vs one call graphs={graphs}")
    cfg = model_h(32, seed=9)
    st = state_for(cfg, 31)
    with shhd.Batch(use_graphs=graphs) as one, shhd.Batch(use_graphs=graphs) as split:
        a = one.admit(cfg, st)
        b = split.admit(cfg, st)
        for bb, slot in ((one, a), (split, b)):
            bb.set_flux_request(slot, density=True, order_parameter=True, momentum=True)
        one.step(14)
        for chunk in (1, 1, 4, 2, 1, 5):
            split.step(chunk)
        assert np.array_equal(one.get_state(a), split.get_state(b))
        assert np.array_equal(one.get_flux(a), split.get_flux(b))


def test_admission_budget_and_retire():
    banner("admission by footprint")
    cfg = quiescent(64, 1)
    fp = shhd.Batch.footprint(cfg)
    assert fp > 0
    with shhd.Batch(budget_bytes=int(2.5 * fp)) as b:
        s0 = b.admit(cfg, state_for(cfg, 1))
        s1 = b.admit(cfg, state_for(cfg, 2))
        with pytest.raises(shhd.ShhdError) as ei:
            b.admit(cfg, state_for(cfg, 3))
        assert ei.value.status == shhd.ERR_FULL
        b.step(3)
        b.retire(s0)
        s2 = b.admit(cfg, state_for(cfg, 3))
        b.step(3)
        assert b.usage()[0] == 2
        assert np.all(np.isfinite(b.get_state(s1))) and np.all(np.isfinite(b.get_state(s2)))


@pytest.mark.parametrize("graphs", [False, True])
def test_held_instance_keeps_its_state_and_resumes(graphs):
    """A held instance takes no step (state, step counter and noise stream
    unchanged) while the others advance; released, it continues exactly as if
    it had never been held."""
    banner(f"held instance resumes, graphs={graphs}")
    ca, cb = model_h(32, seed=7), quiescent(64, 104)
    sa, sb = state_for(ca, 11), state_for(cb, 12)
    ref_a = alone(ca, sa, 8, graphs)
    ref_b = alone(cb, sb, 13, graphs)
    with shhd.Batch(use_graphs=graphs) as b:
        ia, ib = b.admit(ca, sa), b.admit(cb, sb)
        b.step(3)
        b.hold(ia)
        b.step(5)                       # only b advances
        b.hold(ia, False)
        b.step(5)
        got_a, got_b = b.get_state(ia), b.get_state(ib)
    assert np.array_equal(got_a, ref_a)
    assert np.array_equal(got_b, ref_b)
