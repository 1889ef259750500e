"""Closed-form pins of the oracle's nonlinear compressible terms.

The products the right-hand side evaluates on the padded grid are
  v = j / rho                        (src/fcalculator_workspace.cc:122-133)
  Pi = j (x) v                       (src/fcalculator.cc:400-494)
  psi v                              (src/fcalculator.cc:226-239, 260-273)
  compressible viscous stress        (src/fcalculator.cc:343-398)
  -A v_dir  (force/gradient)         (src/fcalculator.cc:776-840)
Each test below builds a state for which the mathematics fixes the answer
without the oracle's machinery:

* quadratic exactness: with a uniform density rho0 every product is quadratic
  in the spectra, and a quadratic product of active modes is alias-free under
  the 3/2 rule, so the right-hand side at every active mode equals the exact
  spectral convolution.  The test evaluates it with numpy's FFT on a 2x
  zero-padded grid (also alias-free) and applies -i k . (.) itself.
* Galilean boost: rho = rho0 (1 + eps cos k1.x) and j = rho U make v = U at
  every grid point, so Pi = rho U (x) U, psi v = psi U and the viscous stress
  vanishes (v has no k != 0 mode): RHS_j = -i k c^2 rho_hat - i (k.U) U rho_hat,
  RHS_psi = linear - i (k.U) psi_hat, RHS_rho = -i (k.U) rho_hat.
* single modes: a longitudinal momentum wave decays as -(eta + zeta) k^2 j/rho0,
  a transverse one as -eta k^2 j/rho0 (the reference's stress
  -i(2 eta k v - (eta - zeta)(k.v) 1) gives RHS = -eta k^2 v - zeta k (k.v)).

A wrong component, sign, density (j * rho, j / rho0 in compressible mode) or
viscous combination fails at least one of them.
"""
import math

import numpy as np
import pytest

import shhd
import oracle_py as op


def _cfg(mode=shhd.MODE_COMPRESSIBLE, n=2, eta=0.37, zeta=0.23, c2=0.81, adv=True, grad=(), na=(16, 12),
         lengths=(11.0, 7.5)):
    return shhd.SolverConfig(
        nx_active=na[0], ny_active=na[1], dealias="3/2", lx=lengths[0], ly=lengths[1], num_order_parameters=n,
        mode=mode, integrator=shhd.INTEGRATOR_SRK3, dt=0.01, pressure_coefficient=c2 if mode == 0 else 0.0,
        eta=eta, zeta=zeta if mode == 0 else 0.0, mobility=[0.7 + 0.2 * a for a in range(n)],
        free_energy_kind=shhd.FREE_ENERGY_PHI4 if n else shhd.FREE_ENERGY_NULL,
        mu_k0=[-0.6 + 0.1 * a for a in range(n)], mu_k2=[0.9] * n, mu_k4=[0.05] * n, phi4_u=[0.0] * n,
        momentum_advection=adv and mode != shhd.MODE_QUIESCENT,
        order_parameter_advection=adv and n > 0 and mode != shhd.MODE_QUIESCENT,
        gradient_forces=list(grad))


def _random_field(cfg, rng, amp):
    mask = shhd.active_mask(cfg)
    z = (rng.normal(size=mask.shape) + 1j * rng.normal(size=mask.shape)) * mask
    st = np.zeros(cfg.state_shape(), complex)
    st[0] = z
    st = shhd.hermitian_project(st, cfg)
    f = st[0]
    f[0, 0] = 0.0
    m2 = cfg.nx * cfg.ny
    rms = math.sqrt(2.0 * np.sum(np.abs(f) ** 2)) / m2
    return f * (amp / rms)


def _to_phys(cfg, f, py, px):
    """Band-limited field of the M-grid spectrum f (unscaled forward
    transform, src/fourier_transform.h:43), sampled on a py x px grid."""
    big = np.zeros((py, px // 2 + 1), complex)
    _, _, sy = shhd.wavenumbers(cfg)
    for gy in range(cfg.ny):
        big[sy[gy] % py, : cfg.nx // 2 + 1] = f[gy]
    return np.fft.irfft2(big * (py * px) / (cfg.nx * cfg.ny), s=(py, px))


def _to_spec(cfg, p):
    """Unscaled forward transform of a py x px field, expressed on the M grid
    and restricted to its half-spectrum (the 2x grid carries every mode the
    M grid's active set can hold)."""
    py, px = p.shape
    big = np.fft.rfft2(p) * (cfg.nx * cfg.ny) / (py * px)
    _, _, sy = shhd.wavenumbers(cfg)
    out = np.zeros((cfg.ny, cfg.nx // 2 + 1), complex)
    for gy in range(cfg.ny):
        out[gy] = big[sy[gy] % py, : cfg.nx // 2 + 1]
    return out


def _grid2(cfg):
    return 2 * cfg.ny_active, 2 * cfg.nx_active


def _div(kx, ky, fx, fy):
    return -1j * (kx * fx + ky * fy)


def _linear_psi(cfg, psi, a, kx, ky):
    k2 = kx ** 2 + ky ** 2
    mu = (cfg.mu_k0[a] + cfg.mu_k2[a] * k2 + cfg.mu_k4[a] * k2 * k2) * psi
    return -cfg.mobility[a] * k2 * mu


def _compare(got, want, mask, tol=1e-12):
    scale = max(np.abs(want[mask]).max(), 1e-300)
    err = np.abs(got[mask] - want[mask]).max() / scale
    assert err <= tol, f"max rel deviation {err:.3e}"
    assert np.abs(got[~mask]).max() == 0.0


def _quadratic_state(cfg, rng, rho0):
    n = cfg.num_order_parameters
    st = np.zeros(cfg.state_shape(), complex)
    st[0, 0, 0] = rho0 * cfg.nx * cfg.ny
    for a in range(n):
        st[1 + a] = _random_field(cfg, rng, 0.3 + 0.1 * a)
        st[1 + a][0, 0] = (0.2 - 0.1 * a) * cfg.nx * cfg.ny
    st[n + 1] = _random_field(cfg, rng, 0.05)
    st[n + 2] = _random_field(cfg, rng, 0.04)
    st[n + 1][0, 0] = 0.011 * cfg.nx * cfg.ny
    st[n + 2][0, 0] = -0.007 * cfg.nx * cfg.ny
    return st


@pytest.mark.parametrize("mode", [shhd.MODE_COMPRESSIBLE, shhd.MODE_INCOMPRESSIBLE])
def test_uniform_density_quadratic_terms_are_exact_convolutions(mode):
    cfg = _cfg(mode=mode)
    rng = np.random.default_rng(11 + mode)
    rho0 = 1.37
    st = _quadratic_state(cfg, rng, rho0)
    n = cfg.num_order_parameters
    got = op.Oracle(cfg).rhs_det(st)

    kx, ky, _ = shhd.wavenumbers(cfg)
    mask = shhd.active_mask(cfg)
    py, px = _grid2(cfg)
    jx, jy = _to_phys(cfg, st[n + 1], py, px), _to_phys(cfg, st[n + 2], py, px)
    vx, vy = jx / rho0, jy / rho0
    want = np.zeros_like(got)
    if mode == shhd.MODE_COMPRESSIBLE:
        want[0] = np.where(mask, _div(kx, ky, st[n + 1], st[n + 2]), 0)
    for a in range(n):
        psi = _to_phys(cfg, st[1 + a], py, px)
        adv = _div(kx, ky, _to_spec(cfg, psi * vx), _to_spec(cfg, psi * vy))
        want[1 + a] = np.where(mask, _linear_psi(cfg, st[1 + a], a, kx, ky) + adv, 0)
    pxx, pxy, pyy = _to_spec(cfg, jx * vx), _to_spec(cfg, jx * vy), _to_spec(cfg, jy * vy)
    vhx, vhy = st[n + 1] / rho0, st[n + 2] / rho0
    kv = kx * vhx + ky * vhy
    k2 = kx ** 2 + ky ** 2
    if mode == shhd.MODE_COMPRESSIBLE:   # -eta k^2 v - zeta k (k.v)
        visc_x = -cfg.eta * k2 * vhx - cfg.zeta * kx * kv
        visc_y = -cfg.eta * k2 * vhy - cfg.zeta * ky * kv
    else:                                # -eta (k^2 v + k (k.v)): symmetric gradient, src/fcalculator.cc:302-341
        visc_x = -cfg.eta * (k2 * vhx + kx * kv)
        visc_y = -cfg.eta * (k2 * vhy + ky * kv)
    want[n + 1] = np.where(mask, visc_x + _div(kx, ky, pxx, pxy), 0)
    want[n + 2] = np.where(mask, visc_y + _div(kx, ky, pxy, pyy), 0)
    for f in range(cfg.num_fields):
        _compare(got[f], want[f], mask)


def test_galilean_boost_makes_velocity_uniform():
    cfg = _cfg()
    n = cfg.num_order_parameters
    rng = np.random.default_rng(5)
    rho0, eps, U = 1.3, 0.3, (0.37, -0.21)
    m2 = cfg.nx * cfg.ny
    st = np.zeros(cfg.state_shape(), complex)
    rho = np.zeros((cfg.ny, cfg.nx // 2 + 1), complex)
    rho[0, 0] = rho0 * m2
    g1x, g1y = 2, 1   # rho0 eps cos(k1 . x): half amplitude at (g1x, g1y) in the half spectrum
    rho[g1y, g1x] = 0.5 * eps * rho0 * m2
    rho[cfg.ny - 3, 1] = 0.25 * eps * rho0 * m2 * (0.6 - 0.8j)   # a second, complex mode
    st[0] = rho
    st[n + 1] = U[0] * rho
    st[n + 2] = U[1] * rho
    for a in range(n):
        st[1 + a] = _random_field(cfg, rng, 0.3)
    got = op.Oracle(cfg).rhs_det(st)

    kx, ky, _ = shhd.wavenumbers(cfg)
    mask = shhd.active_mask(cfg)
    kU = kx * U[0] + ky * U[1]
    want = np.zeros_like(got)
    want[0] = np.where(mask, -1j * kU * rho, 0)
    for a in range(n):
        want[1 + a] = np.where(mask, _linear_psi(cfg, st[1 + a], a, kx, ky) - 1j * kU * st[1 + a], 0)
    c2 = cfg.pressure_coefficient
    want[n + 1] = np.where(mask, -1j * kx * c2 * rho - 1j * kU * U[0] * rho, 0)
    want[n + 2] = np.where(mask, -1j * ky * c2 * rho - 1j * kU * U[1] * rho, 0)
    for f in range(cfg.num_fields):
        _compare(got[f], want[f], mask, tol=1e-11)


@pytest.mark.parametrize("polarization,decay", [("longitudinal", "eta+zeta"), ("transverse", "eta")])
def test_single_momentum_mode_viscous_decay(polarization, decay):
    cfg = _cfg(n=0, adv=False, c2=0.0)
    rho0 = 0.8
    m2 = cfg.nx * cfg.ny
    gx, gy = 3, 2
    kx, ky, _ = shhd.wavenumbers(cfg)
    k = np.array([kx[0, gx], ky[gy, 0]])
    e = k / np.linalg.norm(k) if polarization == "longitudinal" else np.array([-k[1], k[0]]) / np.linalg.norm(k)
    amp = (0.02 - 0.013j) * m2
    st = np.zeros(cfg.state_shape(), complex)
    st[0, 0, 0] = rho0 * m2
    st[1, gy, gx] = e[0] * amp
    st[2, gy, gx] = e[1] * amp
    got = op.Oracle(cfg).rhs_det(st)
    rate = cfg.eta + cfg.zeta if decay == "eta+zeta" else cfg.eta
    k2 = float(k @ k)
    for f in (1, 2):
        want = -rate * k2 * st[f, gy, gx] / rho0
        assert abs(got[f, gy, gx] - want) <= 1e-12 * abs(amp) * k2
        rest = got[f].copy()
        rest[gy, gx] = 0
        assert np.abs(rest).max() <= 1e-12 * abs(amp)


@pytest.mark.parametrize("mode", [shhd.MODE_COMPRESSIBLE, shhd.MODE_INCOMPRESSIBLE])
def test_gradient_force_is_minus_amplitude_times_velocity(mode):
    grad = [dict(component=0, direction=0, amplitude=0.45), dict(component=1, direction=1, amplitude=-0.3),
            dict(component=1, direction=0, amplitude=0.2)]
    cfg = _cfg(mode=mode, adv=False, eta=0.0, zeta=0.0, grad=grad)
    rng = np.random.default_rng(21 + mode)
    rho0 = 1.25
    st = _quadratic_state(cfg, rng, rho0)
    n = cfg.num_order_parameters
    got = op.Oracle(cfg).rhs_det(st)
    kx, ky, _ = shhd.wavenumbers(cfg)
    mask = shhd.active_mask(cfg)
    for a in range(n):
        src = np.zeros_like(st[0])
        for g in grad:
            if g["component"] == a:
                src = src - g["amplitude"] * st[n + 1 + g["direction"]] / rho0
        want = np.where(mask, _linear_psi(cfg, st[1 + a], a, kx, ky) + src, 0)
        _compare(got[1 + a], want, mask)


def test_gradient_force_under_boost_is_a_uniform_source():
    """v = U everywhere: -A U_x is a constant, whose transform is -A U_x Mx My
    at k = 0 only (the force adds every active mode, k = 0 included)."""
    grad = [dict(component=0, direction=0, amplitude=0.45)]
    cfg = _cfg(n=1, adv=False, eta=0.0, zeta=0.0, grad=grad)
    m2 = cfg.nx * cfg.ny
    rho0, U = 1.1, (0.3, 0.2)
    st = np.zeros(cfg.state_shape(), complex)
    st[0, 0, 0] = rho0 * m2
    st[0, 2, 1] = 0.1 * rho0 * m2
    st[0] = shhd.hermitian_project(st, cfg)[0]
    st[2] = U[0] * st[0]
    st[3] = U[1] * st[0]
    got = op.Oracle(cfg).rhs_det(st)
    want = np.zeros_like(got[1])
    want[0, 0] = -0.45 * U[0] * m2
    assert np.abs(got[1] - want).max() <= 1e-12 * m2
