"""FluxBuffer recording of the oracle (src/buffer_flux.{h,cc}, the flux
records of src/fcalculator.cc and the accumulate calls of
src/time_integrator_*.h), pinned by what the conservation laws fix:

* every right-hand-side term of rho, psi_a and j is the divergence of a
  recorded flux, and the SRK3 stage weights (1/6, 1/6, 2/3) and the
  stochastic draw A scaled by 1/sqrt(dt) are exactly the combination that
  the Shu-Osher stages apply to the state (the B draw cancels:
  b1/6 + b2/6 + 2 b3/3 = 0), so  u(t+dt) - u(t) = dt (-i k . F)  holds mode
  by mode for every integrator and mode, noise included; a wrong weight,
  slot, sign, component or noise scaling breaks it;
* incompressible: the projection pressure added to Pi_xx and Pi_yy makes
  -i k . Pi the projected increment that the integrator applies;
* the k = 0 mode, which the divergence does not see: a fluid at rest has
  Pi_xx = Pi_yy = c_T^2 rho_hat(0) and nothing else;
* request semantics: slots that are not requested stay zero.
"""
import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import flux_configs, flux_divergence, flux_state

CONFIGS = flux_configs()


def _evolving(cfg):
    n = cfg.num_order_parameters
    if cfg.mode == shhd.MODE_COMPRESSIBLE:
        return list(range(n + 3))
    if cfg.mode == shhd.MODE_INCOMPRESSIBLE:
        return list(range(1, n + 3))
    return list(range(1, n + 1))


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_flux_divergence_is_the_step_increment(name):
    cfg = CONFIGS[name]
    o = op.Oracle(cfg)
    o.set_state(flux_state(cfg))
    o.set_flux_request(density=True, order_parameter=True, momentum=True)
    o.step(2)
    before = o.get_state()
    o.step(1, t0=2 * cfg.dt)
    after = o.get_state()
    flux = o.get_flux()
    div = cfg.dt * flux_divergence(cfg, flux)
    inc = after - before
    for f in range(cfg.num_fields):
        scale = np.abs(inc[f]).max()
        if f in _evolving(cfg):
            assert scale > 0
            assert np.abs(inc[f] - div[f]).max() <= 1e-11 * scale, (name, f)
        else:
            assert scale == 0
    # the flux lives on the active modes only
    assert np.all(flux[:, ~shhd.active_mask(cfg)] == 0)


def test_rest_state_pressure_at_k0():
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.01,
                            pressure_coefficient=1.7, eta=0.3, zeta=0.2, momentum_advection=True)
    st = np.zeros(cfg.state_shape(), complex)
    st[0, 0, 0] = 1.3 * cfg.nx * cfg.ny
    o = op.Oracle(cfg)
    o.set_state(st)
    o.set_flux_request(density=True, momentum=True)
    o.step(1)
    flux = o.get_flux()
    p = 1.7 * 1.3 * cfg.nx * cfg.ny
    assert abs(flux[2, 0, 0] - p) <= 1e-12 * p      # Pi_xx
    assert abs(flux[4, 0, 0] - p) <= 1e-12 * p      # Pi_yy
    flux[2, 0, 0] = flux[4, 0, 0] = 0
    assert np.abs(flux).max() <= 1e-12 * p


def test_density_flux_is_momentum():
    """rho_det records j_hat itself: with Euler the density flux of a step is
    the momentum at its start (src/fcalculator.cc:60-84)."""
    cfg = CONFIGS["model_h_euler"]
    o = op.Oracle(cfg)
    st = flux_state(cfg)
    o.set_state(st)
    o.set_flux_request(density=True)
    o.step(1)
    flux = o.get_flux()
    n = cfg.num_order_parameters
    assert np.array_equal(flux[0], st[n + 1])
    assert np.array_equal(flux[1], st[n + 2])
    assert np.all(flux[2:] == 0)


@pytest.mark.parametrize("req", [(1, 0, 0), (0, 1, 0), (0, 0, 1)])
def test_only_requested_slots_record(req):
    cfg = CONFIGS["model_h_srk3"]
    o = op.Oracle(cfg)
    o.set_state(flux_state(cfg))
    o.set_flux_request(*req)
    o.step(1)
    flux = o.get_flux()
    n = cfg.num_order_parameters
    groups = [slice(0, 2), slice(2, 2 + 2 * n), slice(2 + 2 * n, 5 + 2 * n)]
    for on, sl in zip(req, groups):
        if on:
            assert np.abs(flux[sl]).max() > 0
        else:
            assert np.all(flux[sl] == 0)
