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
