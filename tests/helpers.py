"""State builders and comparison helpers shared by the tests."""
import math

import numpy as np

import shhd


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    den = np.linalg.norm(b.ravel())
    return float(np.linalg.norm((a - b).ravel()) / (den if den > 0 else 1.0))


def random_active_state(cfg, rng, amp=None, rho0=1.0):
    """Random Hermitian, masked spectral state with uniform density rho0 plus
    small fluctuations (amplitudes in physical units of `amp` per field)."""
    nf = cfg.num_fields
    m2 = cfg.nx * cfg.ny
    amp = amp or {}
    st = np.zeros(cfg.state_shape(), complex)
    mask = shhd.active_mask(cfg)
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    env = np.exp(-0.5 * k2 / (k2.max() * 0.15 + 1e-30))
    for f in range(nf):
        a = amp.get(f, 0.0)
        if a == 0.0:
            continue
        z = (rng.normal(size=mask.shape) + 1j * rng.normal(size=mask.shape)) * env
        st[f] = np.where(mask, z, 0)
    st = shhd.hermitian_project(st, cfg)
    # scale each field to the requested physical rms
    nmodes = mask.sum()
    for f in range(nf):
        a = amp.get(f, 0.0)
        if a == 0.0:
            continue
        rms = math.sqrt(2.0 * np.sum(np.abs(st[f]) ** 2)) / m2
        st[f] *= a / max(rms, 1e-300)
        st[f][0, 0] = 0.0
    st[0][0, 0] += rho0 * m2
    _ = nmodes
    return st


def sine_momentum_state(cfg, amp, base_rho=1.0, nky=1):
    """set density uniform + set momentum x sine nkx 0 nky (src/initial_condition_sine_momentum.cc:30-62)."""
    m2 = cfg.nx * cfg.ny
    st = np.zeros(cfg.state_shape(), complex)
    st[0, 0, 0] = base_rho * m2
    jx = cfg.num_order_parameters + 1
    st[jx, nky, 0] = -0.5j * amp * m2
    st[jx, cfg.ny - nky, 0] = 0.5j * amp * m2
    return st


def flux_divergence(cfg, flux):
    """-i k . F of every conservation law, in state layout [rho, psi_a, jx, jy],
    from the FluxBuffer slots [rho_x, rho_y, (psi_x, psi_y)_a, Pi_xx, Pi_xy, Pi_yy]
    (src/buffer_flux.cc:10-48)."""
    kx, ky, _ = shhd.wavenumbers(cfg)
    n = cfg.num_order_parameters
    out = np.zeros(cfg.state_shape(), complex)
    out[0] = -1j * (kx * flux[0] + ky * flux[1])
    for a in range(n):
        out[1 + a] = -1j * (kx * flux[2 + 2 * a] + ky * flux[3 + 2 * a])
    b = 2 + 2 * n
    out[n + 1] = -1j * (kx * flux[b] + ky * flux[b + 1])
    out[n + 2] = -1j * (kx * flux[b + 1] + ky * flux[b + 2])
    return out


def flux_configs():
    """Configurations whose every right-hand-side term is the divergence of a
    recorded flux (no sine or gradient forces), one per integrator and mode,
    on a non-square grid with unequal lengths."""
    base = dict(nx_active=32, ny_active=16, dealias="3/2", lx=20.0, ly=13.0)
    phi4 = lambda n: dict(num_order_parameters=n, mobility=[1.0 + 0.3 * a for a in range(n)],  # noqa: E731
                          free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[-1.0] * n, mu_k2=[0.2] * n,
                          mu_k4=[0.002] * n, phi4_u=[1.0] * n)
    out = {}
    for integ, tag in ((shhd.INTEGRATOR_SRK3, "srk3"), (shhd.INTEGRATOR_EULER, "euler")):
        out[f"model_h_{tag}"] = shhd.SolverConfig(
            **base, **phi4(2), integrator=integ, dt=0.004, pressure_coefficient=1.0, eta=0.5, zeta=0.3,
            momentum_advection=True, order_parameter_advection=True, noise_momentum=True,
            noise_order_parameter=True, kBT=1e-4)
        out[f"incompressible_{tag}"] = shhd.SolverConfig(
            **base, **phi4(1), integrator=integ, dt=0.004, mode=shhd.MODE_INCOMPRESSIBLE, eta=0.4,
            momentum_advection=True, order_parameter_advection=True, noise_momentum=True,
            noise_order_parameter=True, kBT=1e-4)
        out[f"quiescent_{tag}"] = shhd.SolverConfig(
            **base, **phi4(1), integrator=integ, dt=0.004, mode=shhd.MODE_QUIESCENT,
            noise_order_parameter=True, kBT=0.02)
    # three order parameters: the flux slots of an order parameter beyond the
    # first two, through the row pass built for up to four
    out["model_h3_srk3"] = shhd.SolverConfig(
        **base, **phi4(3), dt=0.004, pressure_coefficient=1.0, eta=0.5, zeta=0.3, momentum_advection=True,
        order_parameter_advection=True, noise_momentum=True, noise_order_parameter=True, kBT=1e-4)
    return out


def flux_state(cfg, seed=3):
    rng = np.random.default_rng(seed)
    n = cfg.num_order_parameters
    amp = {0: 0.01, n + 1: 0.02, n + 2: 0.02}
    for a in range(n):
        amp[1 + a] = 0.3
    st = random_active_state(cfg, rng, amp=amp)
    if n:
        st[1, 0, 0] = 0.1 * cfg.nx * cfg.ny
    return st
