"""Device-side initial conditions (shhd_*_apply_initial_condition) against the
oracle's restatement of src/initial_condition_*.cc: FP64 fields agree to
1e-13 relative (only libm rounding of log/sqrt/cos/sin/exp differs; the
splitmix64 hash is integer-exact), FP32 to 1e-6; the same zero pattern."""
import math

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import rel_l2

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[case] {name}", flush=True)


def cfg_for(na, n, mode, precision, lx=None, ly=None, dealias="3/2"):
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias=dealias, lx=lx or 0.8 * na, ly=ly or 1.1 * na,
                             dt=0.005, mode=mode, num_order_parameters=n, mobility=[1.0] * n,
                             free_energy_kind=shhd.FREE_ENERGY_PHI4 if n else shhd.FREE_ENERGY_NULL,
                             mu_k0=[0.5] * n, mu_k2=[1.0] * n, phi4_u=[0.2] * n, eta=0.3, zeta=0.3,
                             pressure_coefficient=1.0 if mode == shhd.MODE_COMPRESSIBLE else 0.0,
                             momentum_advection=mode != shhd.MODE_QUIESCENT, precision=precision)


def ic_sets(cfg):
    D = shhd.initial_condition(shhd.IC_EQUILIBRIUM_DENSITY, seed=5, kBT=0.01, value=1.2, pressure_coefficient=1.0)
    D0 = shhd.initial_condition(shhd.IC_EQUILIBRIUM_DENSITY, seed=5, kBT=0.0, value=1.0, pressure_coefficient=1.0)
    P = [shhd.initial_condition(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, component=a, seed=9, kBT=0.02, value=0.1 * a,
                                k0=0.5, k2=1.0) for a in range(cfg.num_order_parameters)]
    out = {}
    if cfg.mode == shhd.MODE_COMPRESSIBLE:
        out["equilibrium_compressible"] = [D] + P + [
            shhd.initial_condition(shhd.IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE, component=d, seed=2, kBT=0.01)
            for d in (0, 1)]
    if cfg.mode == shhd.MODE_INCOMPRESSIBLE:
        out["equilibrium_incompressible"] = [D0] + P + [
            shhd.initial_condition(shhd.IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE, component=d, seed=2, kBT=0.01)
            for d in (0, 1)]
        out["random_vorticity"] = [D0] + [shhd.initial_condition(shhd.IC_RANDOM_VORTICITY, component=d, seed=3,
                                                                 n0=8.0, sigma=2.0, urms=1.0) for d in (0, 1)]
        out["spectrum_k4"] = [D0] + [shhd.initial_condition(shhd.IC_SPECTRUM_K4, component=d, seed=3, n0=8.0,
                                                            urms=1.0) for d in (0, 1)]
    if cfg.mode == shhd.MODE_QUIESCENT:
        out["equilibrium_order_parameter"] = [D0] + P
    return out


CASES = [(64, 2, shhd.MODE_COMPRESSIBLE), (128, 0, shhd.MODE_INCOMPRESSIBLE), (64, 1, shhd.MODE_QUIESCENT)]


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("na,n,mode", CASES)
def test_device_ic_matches_oracle(na, n, mode, precision):
    cfg = cfg_for(na, n, mode, precision)
    for name, ics in ic_sets(cfg).items():
        banner(f"ic {name} {na}^2 N={n} fp{precision}")
        o = op.Oracle(cfg)
        o.set_state(np.zeros(cfg.state_shape(), complex))
        with shhd.GpuFluidSolver(cfg) as g:
            for ic in ics:
                o.apply_initial_condition(ic)
                g.apply_initial_condition(ic)
            got = g.get_state()
        ref = o.get_state()
        assert np.array_equal(got == 0, ref == 0) or precision == 32
        tol = 1e-13 if precision == 64 else 1e-6
        for f in range(cfg.num_fields):
            if np.linalg.norm(ref[f]) == 0:
                assert np.all(got[f] == 0)
                continue
            e = rel_l2(got[f], ref[f])
            assert e <= tol, (name, f, e)
            if precision == 64:
                assert np.abs(got[f] - ref[f]).max() <= 1e-13 * np.abs(ref[f]).max()


def test_ic_then_steps_match_oracle():
    """A state built on the device then advanced equals the oracle's run from
    its own initial conditions (FP64, 20 SRK3 steps, <= 1e-10)."""
    cfg = cfg_for(64, 2, shhd.MODE_COMPRESSIBLE, 64)
    ics = ic_sets(cfg)["equilibrium_compressible"]
    banner("ic + 20 steps")
    o = op.Oracle(cfg)
    o.set_state(np.zeros(cfg.state_shape(), complex))
    with shhd.GpuFluidSolver(cfg) as g:
        for ic in ics:
            o.apply_initial_condition(ic)
            g.apply_initial_condition(ic)
        o.step(20)
        g.step(20)
        got = g.get_state()
    ref = o.get_state()
    for f in range(cfg.num_fields):
        assert rel_l2(got[f], ref[f]) <= 1e-10


def test_ic_errors_and_batch():
    cfg = cfg_for(32, 1, shhd.MODE_INCOMPRESSIBLE, 64)
    banner("ic errors / batch")
    with shhd.GpuFluidSolver(cfg) as g:
        with pytest.raises(shhd.ShhdError, match="requires positive reference density"):
            g.apply_initial_condition(shhd.initial_condition(shhd.IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE, kBT=0.1))
        with pytest.raises(shhd.ShhdError, match="positive linear chemical-potential"):
            g.apply_initial_condition(shhd.initial_condition(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, kBT=0.1, k0=-1.0))
        with pytest.raises(shhd.ShhdError, match="requires positive sigma"):
            g.apply_initial_condition(shhd.initial_condition(shhd.IC_RANDOM_VORTICITY, n0=4.0, urms=1.0))
    ics = ic_sets(cfg)["spectrum_k4"]
    with shhd.GpuFluidSolver(cfg) as g:
        for ic in ics:
            g.apply_initial_condition(ic)
        single = g.get_state()
    with shhd.Batch() as b:
        slot = b.admit(cfg, np.zeros(cfg.state_shape(), complex) + single * 0)
        for ic in ics:
            b.apply_initial_condition(slot, ic)
        assert np.array_equal(b.get_state(slot), single)
