"""Flux recording of the HIP path (shhd_gpu_set_flux_request / shhd_gpu_get_flux
and the batch variants; FluxBuffer of src/buffer_flux.{h,cc}).

* Deterministic runs (noise off): every flux slot agrees with the oracle's to
  <= 1e-10 relative L2 in FP64 and <= 1e-5 in FP32; slots the oracle leaves
  at zero are zero.
* Noisy runs: the device's own fluxes satisfy u(t+dt) - u(t) = dt (-i k . F)
  field by field (the identity that pins the oracle's flux), to 1e-10 in FP64
  and 1e-3 in FP32 (there the increment is a difference of FP32-stored states).
* Recording does not change the trajectory; batched instances record exactly
  what they record alone; a changed request clears the slots it drops.
"""
import dataclasses

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import flux_configs, flux_divergence, flux_state, rel_l2

pytestmark = pytest.mark.gpu

CONFIGS = flux_configs()


def banner(name):
    print(f"[case] {name}", flush=True)


def noiseless(cfg, precision=64):
    return dataclasses.replace(cfg, noise_momentum=False, noise_order_parameter=False, precision=precision)


def evolving(cfg):
    n = cfg.num_order_parameters
    if cfg.mode == shhd.MODE_COMPRESSIBLE:
        return list(range(n + 3))
    if cfg.mode == shhd.MODE_INCOMPRESSIBLE:
        return list(range(1, n + 3))
    return list(range(1, n + 1))


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_flux_matches_oracle(name, precision):
    cfg = noiseless(CONFIGS[name], precision)
    banner(f"flux parity {name} fp{precision}")
    tol = 1e-10 if precision == 64 else 1e-5
    st = flux_state(cfg)
    o = op.Oracle(cfg)
    o.set_state(st)
    o.set_flux_request(True, True, True)
    o.step(3)
    ref = o.get_flux()
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        g.set_flux_request(True, True, True)
        g.step(3)
        got = g.get_flux()
    errs = []
    for k in range(ref.shape[0]):
        if np.linalg.norm(ref[k]) == 0:
            assert np.all(got[k] == 0), f"slot {k} should stay zero"
            errs.append(0.0)
            continue
        errs.append(rel_l2(got[k], ref[k]))
    print(f"    per-slot rel L2 {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= tol
    assert np.all(got[:, ~shhd.active_mask(cfg)] == 0)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_noisy_flux_divergence_is_the_step_increment(name, precision):
    cfg = dataclasses.replace(CONFIGS[name], precision=precision)
    banner(f"flux identity {name} fp{precision}")
    tol = 1e-10 if precision == 64 else 1e-3
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(flux_state(cfg))
        g.set_flux_request(True, True, True)
        g.step(2)
        before = g.get_state()
        g.step(1, 2 * cfg.dt)
        after = g.get_state()
        flux = g.get_flux()
    inc = after - before
    div = cfg.dt * flux_divergence(cfg, flux)
    errs = []
    for f in range(cfg.num_fields):
        if f in evolving(cfg):
            assert np.abs(inc[f]).max() > 0
            errs.append(rel_l2(div[f], inc[f]))
        else:
            assert np.all(inc[f] == 0) and np.all(div[f] == 0)
    print(f"    per-field rel L2 {['%.1e' % e for e in errs]}", flush=True)
    assert max(errs) <= tol


def test_recording_leaves_the_trajectory_unchanged():
    cfg = CONFIGS["model_h_srk3"]
    banner("flux on/off trajectories")
    st = flux_state(cfg)
    out = []
    for req in (False, True):
        with shhd.GpuFluidSolver(cfg) as g:
            g.set_state(st)
            g.set_flux_request(req, req, req)
            g.step(7)
            out.append(g.get_state())
    assert np.array_equal(out[0], out[1])


@pytest.mark.parametrize("graphs", [False, True])
def test_batch_flux_alone_and_mixed(graphs):
    banner(f"batch flux graphs={graphs}")
    names = ["model_h_srk3", "incompressible_euler", "quiescent_srk3"]
    cfgs = [CONFIGS[n] for n in names]
    states = [flux_state(c, seed=10 + i) for i, c in enumerate(cfgs)]
    reqs = [(True, True, True), (False, True, True), (False, True, False)]
    alone = []
    for c, s, r in zip(cfgs, states, reqs):
        with shhd.Batch(use_graphs=graphs) as b:
            slot = b.admit(c, s)
            b.set_flux_request(slot, *r)
            b.step(7)
            alone.append(b.get_flux(slot))
    with shhd.Batch(use_graphs=graphs) as b:
        slots = [b.admit(c, s) for c, s in zip(cfgs, states)]
        for slot, r in zip(slots, reqs):
            b.set_flux_request(slot, *r)
        b.step(7)
        for i, slot in enumerate(slots):
            assert np.array_equal(b.get_flux(slot), alone[i]), names[i]
        # drop the momentum request of instance 0: its Pi slots are cleared at
        # the next step, the others still record
        b.set_flux_request(slots[0], True, True, False)
        b.step(3)
        f0 = b.get_flux(slots[0])
        n = cfgs[0].num_order_parameters
        assert np.all(f0[2 + 2 * n:] == 0)
        assert np.abs(f0[:2 + 2 * n]).max() > 0


def test_flux_size_is_checked():
    cfg = CONFIGS["quiescent_euler"]
    banner("flux size check")
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(flux_state(cfg))
        g.set_flux_request(order_parameter=True)
        g.step(1)
        bad = np.zeros(10, np.float64)
        st = g.lib.shhd_gpu_get_flux(g.h, bad.ctypes.data_as(shhd.ctypes.POINTER(shhd.ctypes.c_double)), bad.size)
        assert st == shhd.ERR_INVALID
        f = g.get_flux()
        assert np.abs(f[2:4]).max() > 0
        assert np.all(f[:2] == 0) and np.all(f[4:] == 0)
