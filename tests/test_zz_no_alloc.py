"""The allocation-free initial-condition entry point gives what the allocating
one gives, bit for bit, on two calls in a row (the second on a smaller grid,
whose fewer partial sums must not read the first call's leftovers)."""
import numpy as np
import pytest

import shhd

pytestmark = pytest.mark.gpu


def incompressible(na, precision):
    return shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", dt=1e-4, mode=shhd.MODE_INCOMPRESSIBLE,
                             eta=1e-4, momentum_advection=True, precision=precision)


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("kind", [shhd.IC_RANDOM_VORTICITY, shhd.IC_SPECTRUM_K4])
def test_noalloc_initial_condition_equals_allocating_one(kind, precision):
    print(f"[case] no-alloc initial condition kind {kind} fp{precision}: 2048^2 then 64^2", flush=True)
    IC = shhd.initial_condition
    density = IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=3, kBT=0.0, value=1.3, pressure_coefficient=1.0)
    with shhd.Batch() as b:
        for na in (2048, 64):   # 2048 partial sums, then 8
            cfg = incompressible(na, precision)
            sa = b.admit(cfg)
            sb = b.admit(cfg)
            for s in (sa, sb):
                b.apply_initial_condition(s, density)
            for d in (0, 1):
                ic = IC(kind, component=d, seed=11 + na, n0=6.0, sigma=2.0, urms=0.7)
                b.apply_initial_condition(sa, ic)
                b.apply_initial_condition_noalloc(sb, ic)
            want = b.get_state(sa)
            got = b.get_state(sb)
            assert np.any(want[cfg.num_order_parameters + 1] != 0)
            assert np.array_equal(got, want)
            b.retire(sa)
            b.retire(sb)
