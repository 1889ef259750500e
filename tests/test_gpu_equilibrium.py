"""Hydrodynamic fluctuations in distribution: equipartition and flat S_rho_rho.

The reference's own known answers (examples/01_best_timestep_equilibrium/
README.md:40-53; theory in its plot_results.py: E = kBT/2 per real degree of
freedom, (Na-1)^2 - 1 of them per scalar field): on a 64^2 grid at kBT = 1,
  compressible:   E_K = 3968, E_C = 1984, E_psi = 1984
  incompressible: E_K = 1984,             E_psi = 1984
with the energies of TimeSeriesMeasure (src/measure_time_series.cc:363-425):
  E_K   = 1/2 dA / (Nc rho0) sum_k w |j_hat|^2      (incompressible form; the
          compressible one integrates j^2 / rho, equal to it to O(delta rho))
  E_C   = 1/2 cT^2 dA / (Nc rho0) sum_{k != 0} w |rho_hat|^2
  E_psi = 1/2 a dA / Nc sum_{k != 0} w |psi_hat|^2   (quadratic free energy)
with dA = V / Nc, Nc = Mx My and w = 1 on the kx = 0 column, 2 elsewhere.
The runs use the example scripts' parameters (transport 8.0: eta = zeta =
M = 8, cT = 10, a = 1, nonlinear on, noise on), start from the reference's
equilibrium/gaussian initial conditions and average the device's FP64
structure-factor sums over the run.  Sampling error of one snapshot of E_K
is sqrt(E_K kBT) ~ 1 %; the averages of a few hundred correlated samples are
held to 3 %.

The 128^2 compressible Navier-Stokes configuration (cT = 1, eta = zeta = 0.1, kBT = 1e-4,
dt = 0.05, 2000 steps) must keep a flat S_rho_rho(k) = rho0 kBT / cT^2.
"""
import numpy as np
import pytest

import shhd

pytestmark = pytest.mark.gpu


def banner(name):
    print(f"[case] {name}", flush=True)


def weights(cfg):
    w = np.where(shhd.active_mask(cfg), 2.0, 0.0)
    w[:, 0] = np.where(shhd.active_mask(cfg)[:, 0], 1.0, 0.0)
    w[0, 0] = 0.0   # k = 0 (the conserved totals) carries no fluctuation energy
    return w


def example_cfg(mode, precision):
    comp = mode == shhd.MODE_COMPRESSIBLE
    return shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=64.0, ly=64.0,
                             dt=0.002 if comp else 0.005, mode=mode, num_order_parameters=1,
                             pressure_coefficient=100.0 if comp else 0.0, eta=8.0, zeta=8.0 if comp else 0.0,
                             mobility=[8.0], free_energy_kind=shhd.FREE_ENERGY_QUADRATIC, mu_k0=[1.0],
                             momentum_advection=True, order_parameter_advection=True, noise_momentum=True,
                             noise_order_parameter=True, kBT=1.0, seed=12345, precision=precision)


def equilibrium_ics(mode, kbt, cT2, seed):
    IC = shhd.initial_condition
    mom = (shhd.IC_EQUILIBRIUM_MOMENTUM_COMPRESSIBLE if mode == shhd.MODE_COMPRESSIBLE
           else shhd.IC_EQUILIBRIUM_MOMENTUM_INCOMPRESSIBLE)
    dens_kbt = kbt if mode == shhd.MODE_COMPRESSIBLE else 0.0
    return ([IC(shhd.IC_EQUILIBRIUM_DENSITY, seed=seed, kBT=dens_kbt, value=1.0, pressure_coefficient=cT2)]
            + [IC(mom, component=d, seed=seed, kBT=kbt) for d in (0, 1)])


@pytest.mark.parametrize("mode", [shhd.MODE_COMPRESSIBLE, shhd.MODE_INCOMPRESSIBLE])
@pytest.mark.parametrize("precision", [64, 32])
def test_equipartition_energies(mode, precision):
    comp = mode == shhd.MODE_COMPRESSIBLE
    banner(f"equipartition 64^2 kBT=1 {'compressible' if comp else 'incompressible'} fp{precision}")
    cfg = example_cfg(mode, precision)
    nc = cfg.nx * cfg.ny
    vol = cfg.lx * cfg.ly
    w = weights(cfg)
    ics = equilibrium_ics(mode, 1.0, cfg.pressure_coefficient or 1.0, seed=7)
    ics.append(shhd.initial_condition(shhd.IC_EQUILIBRIUM_ORDER_PARAMETER, component=0, seed=7, kBT=1.0, k0=1.0))
    burn, run, every = 500, 3000, 10
    with shhd.GpuFluidSolver(cfg) as g:
        for ic in ics:
            g.apply_initial_condition(ic)
        g.step(burn)
        g.set_structure_factor([0, 1, 2, 3], nevery=every)
        g.step(run)
        sk, ns = g.get_structure_factor()
        final = g.get_state()
    assert np.all(np.isfinite(final))
    assert ns == run // every
    s = sk / ns
    rho0 = 1.0
    e_k = 0.5 * vol / (nc * nc * rho0) * np.sum(w * (s[2] + s[3]))
    e_psi = 0.5 * 1.0 * vol / (nc * nc) * np.sum(w * s[1])
    dof = (64 - 1) ** 2 - 1
    want_k = (2 if comp else 1) * 0.5 * dof
    print(f"    E_K {e_k:.1f} (want {want_k:.0f})  E_psi {e_psi:.1f} (want {0.5 * dof:.0f})", flush=True)
    assert abs(e_k / want_k - 1) < 0.03
    assert abs(e_psi / (0.5 * dof) - 1) < 0.03
    if comp:
        e_c = 0.5 * cfg.pressure_coefficient * vol / (nc * nc * rho0) * np.sum(w * s[0])
        print(f"    E_C {e_c:.1f} (want {0.5 * dof:.0f})", flush=True)
        assert abs(e_c / (0.5 * dof) - 1) < 0.03


@pytest.mark.parametrize("precision", [64, 32])
def test_ns128_flat_density_structure_factor(precision):
    banner(f"compressible NS 128^2 S_rho_rho flat, fp{precision}")
    kbt = 1e-4
    cfg = shhd.SolverConfig(nx_active=128, ny_active=128, dealias="3/2", lx=128.0, ly=128.0, dt=0.05,
                            pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True,
                            noise_momentum=True, kBT=kbt, seed=1, precision=precision)
    nc = cfg.nx * cfg.ny
    vol = cfg.lx * cfg.ly
    with shhd.GpuFluidSolver(cfg) as g:
        for ic in equilibrium_ics(shhd.MODE_COMPRESSIBLE, kbt, 1.0, seed=1):
            g.apply_initial_condition(ic)
        g.set_structure_factor([0], nevery=10)
        g.step(2000)
        sk, ns = g.get_structure_factor()
    s = sk[0] / ns * vol / (nc * nc)   # S_rho_rho(k) per active mode
    mask = shhd.active_mask(cfg)
    mask[0, 0] = False
    kx, ky, _ = shhd.wavenumbers(cfg)
    k = np.sqrt(kx ** 2 + ky ** 2)
    want = 1.0 * kbt / 1.0
    vals = s[mask]
    print(f"    S_rho_rho mean {vals.mean():.4e} (want {want:.1e}) over {vals.size} modes, {ns} samples", flush=True)
    assert abs(vals.mean() / want - 1) < 0.03
    # flat: each quarter of the |k| range within sampling error of the same value
    edges = np.quantile(k[mask], [0.0, 0.25, 0.5, 0.75, 1.0])
    for lo, hi in zip(edges[:-1], edges[1:]):
        sel = mask & (k >= lo) & (k <= hi)
        m = s[sel].mean()
        print(f"    |k| in [{lo:.2f}, {hi:.2f}]: {m:.4e}", flush=True)
        assert abs(m / want - 1) < 0.06
