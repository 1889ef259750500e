"""Pins the FP64 CPU oracle against what the mathematics fixes.

The reference ships no unit tests or golden vectors for this path and cannot
be built here (it needs heFFTe and FFTW), so each oracle function is checked
against a closed form, a brute-force evaluation or a known answer from the
reference's own example study (examples/01_best_timestep_equilibrium/README.md).
"""
import math

import numpy as np
import pytest

import shhd
import oracle_py as op
from helpers import rel_l2, random_active_state, sine_momentum_state


def dft(x, sign):
    n = len(x)
    j = np.arange(n)
    w = np.exp(sign * 2j * np.pi * (np.outer(j, j) % n) / n)   # exact angles
    return w @ x


# ---------------------------------------------------------------------------
# transforms (src/fourier_transform.h:39-51)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 12, 18, 24, 48, 96, 192, 384])
def test_fft_matches_dft(n):
    rng = np.random.default_rng(n)
    x = rng.normal(size=n) + 1j * rng.normal(size=n)
    for sign in (-1, 1):
        ref = dft(x, sign)
        got = op.fft_c2c(x, sign)
        assert np.abs(got - ref).max() <= 1e-13 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("my,mx", [(4, 6), (12, 12), (12, 18), (24, 12)])
def test_fft2_r2c_brute_force(my, mx):
    rng = np.random.default_rng(my * 100 + mx)
    p = rng.normal(size=(my, mx))
    ref = np.zeros((my, mx // 2 + 1), complex)
    y = np.arange(my)[:, None]
    x = np.arange(mx)[None, :]
    for ky in range(my):
        for kx in range(mx // 2 + 1):
            ref[ky, kx] = np.sum(p * np.exp(-2j * np.pi * (((kx * x) % mx) / mx + ((ky * y) % my) / my)))
    got = op.fft2_r2c(p)
    assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    # backward carries 1/(Mx*My): round trip is the identity
    assert np.abs(op.fft2_c2r(got, mx) - p).max() <= 1e-13


def test_fft2_single_mode_known_answer():
    mx, my = 18, 12
    y, x = np.mgrid[0:my, 0:mx]
    p = np.cos(2 * np.pi * 3 * x / mx) + 2 * np.sin(2 * np.pi * 2 * y / my)
    s = op.fft2_r2c(p)
    n = mx * my
    assert abs(s[0, 3] - n / 2) < 1e-10
    assert abs(s[2, 0] - (-1j) * n) < 1e-10
    assert abs(s[my - 2, 0] - 1j * n) < 1e-10
    s2 = s.copy()
    s2[0, 3] = s2[2, 0] = s2[my - 2, 0] = 0
    assert np.abs(s2).max() < 1e-10


def test_c2r_ignores_imaginary_dc():
    """The half-spectrum C2R reads gx = 0 as Hermitian: an imaginary part on the
    kx = ky = 0 entry does not reach the real output."""
    mx, my = 12, 12
    rng = np.random.default_rng(3)
    p = rng.normal(size=(my, mx))
    s = op.fft2_r2c(p)
    s[0, 0] += 5j
    assert np.abs(op.fft2_c2r(s, mx) - p).max() < 1e-12


# ---------------------------------------------------------------------------
# 3/2 rule (src/spectral_mask.cc, src/fcalculator.cc:189-274)
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("kmode,dealias", [(2, "3/2"), (3, "3/2"), (2, "none"), (3, "none")])
def test_cubic_term_closed_form(kmode, dealias):
    """psi = A cos(k x): psi^3 = A^3 (3 cos kx + cos 3kx)/4, evaluated on the
    compute grid of M points: the 3k harmonic folds onto |3k mod M| there and
    survives only if that wavenumber is active (3/2 padding removes quadratic
    aliasing, not all cubic aliasing)."""
    na = 8
    cfg = shhd.SolverConfig(nx_active=na, ny_active=na, dealias=dealias, lx=2 * np.pi, ly=2 * np.pi,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[0.7],
                            free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[0.3], mu_k2=[0.2], phi4_u=[1.5])
    o = op.Oracle(cfg)
    n = cfg.nx * cfg.ny
    A = 0.4
    st = np.zeros(cfg.state_shape(), complex)
    st[1, 0, kmode] = A / 2 * n
    rhs = o.rhs_det(st)[1]
    expect = np.zeros_like(rhs)
    mob, a, b, u = 0.7, 0.3, 0.2, 1.5
    # linear part at k
    expect[0, kmode] += -mob * kmode ** 2 * (a + b * kmode ** 2) * (A / 2 * n)
    # u psi^3: 3A^3/4 at k (amplitude of cos), A^3/4 at 3k folded onto the grid
    c1 = u * 3 * A ** 3 / 4
    c3 = u * A ** 3 / 4
    expect[0, kmode] += -mob * kmode ** 2 * (c1 / 2 * n)
    m = cfg.nx
    k3 = (3 * kmode) % m
    if k3 > m // 2:
        k3 = m - k3
    if k3 < na // 2:
        expect[0, k3] += -mob * k3 ** 2 * (c3 / 2 * n)
    assert np.abs(rhs - expect).max() <= 1e-12 * np.abs(expect).max()


# ---------------------------------------------------------------------------
# integrators (src/time_integrator_*.h): exact linear amplification
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("integrator", [shhd.INTEGRATOR_SRK3, shhd.INTEGRATOR_EULER])
def test_quiescent_linear_amplification(integrator):
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.01,
                            integrator=integrator, mode=shhd.MODE_QUIESCENT, num_order_parameters=1,
                            mobility=[0.8], free_energy_kind=shhd.FREE_ENERGY_PHI4, mu_k0=[0.5], mu_k2=[0.3],
                            phi4_u=[0.0])
    rng = np.random.default_rng(1)
    st = random_active_state(cfg, rng, amp={1: 0.5})
    o = op.Oracle(cfg)
    o.set_state(st)
    nsteps = 40
    o.step(nsteps)
    got = o.get_state()[1]
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    z = -0.8 * k2 * (0.5 + 0.3 * k2) * cfg.dt
    R = 1 + z + z * z / 2 + z ** 3 / 6 if integrator == shhd.INTEGRATOR_SRK3 else 1 + z
    expect = st[1] * R ** nsteps
    assert rel_l2(got, expect) < 1e-12


def test_incompressible_viscous_decay():
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", dt=0.01, mode=shhd.MODE_INCOMPRESSIBLE,
                            eta=0.2)
    rng = np.random.default_rng(2)
    st = random_active_state(cfg, rng, amp={1: 0.3, 2: 0.3}, rho0=1.7)
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    with np.errstate(invalid="ignore", divide="ignore"):
        tr = np.where(k2 > 0, (-ky * st[1] + kx * st[2]) / k2, 0)
    st[1] = np.where(k2 > 0, -ky * tr, st[1])
    st[2] = np.where(k2 > 0, kx * tr, st[2])
    st = shhd.hermitian_project(st, cfg)
    o = op.Oracle(cfg)
    o.set_state(st)
    o.step(30)
    got = o.get_state()
    z = -0.2 * k2 * cfg.dt / 1.7
    R = (1 + z + z * z / 2 + z ** 3 / 6) ** 30
    assert rel_l2(got[1:3], st[1:3] * R) < 1e-12
    assert np.array_equal(got[0], st[0])


def test_shear_wave_exact():
    """Compressible-NS kT = 0 workload: j_x = A sin(2 pi y/Ly), rho = 1 is an exact
    nonlinear solution decaying with z = -eta k^2 dt."""
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.05,
                            pressure_coefficient=1.0, eta=0.1, zeta=0.1, momentum_advection=True)
    st = sine_momentum_state(cfg, 1e-3)
    o = op.Oracle(cfg)
    o.set_state(st)
    o.step(100)
    got = o.get_state()
    k = 2 * math.pi / 16
    z = -0.1 * k * k * 0.05
    R = (1 + z + z * z / 2 + z ** 3 / 6) ** 100
    assert abs(got[1, 1, 0] / st[1, 1, 0] - R) < 1e-12
    other = got.copy()
    other[1, 1, 0] = other[1, -1, 0] = 0
    other[0, 0, 0] = 0
    assert np.abs(other).max() < 1e-9 * abs(st[1, 1, 0])


def test_invariants_model_h():
    cfg = shhd.SolverConfig(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.005,
                            num_order_parameters=2, pressure_coefficient=1.0, eta=0.5, zeta=0.5,
                            mobility=[1.0, 1.0], free_energy_kind=shhd.FREE_ENERGY_PHI4,
                            mu_k0=[-1.0, -1.0], mu_k2=[1.0, 1.0], phi4_u=[1.0, 1.0],
                            momentum_advection=True, order_parameter_advection=True)
    rng = np.random.default_rng(4)
    st = random_active_state(cfg, rng, amp={0: 0.01, 1: 0.3, 2: 0.3, 3: 0.02, 4: 0.02})
    st[1, 0, 0] = 0.2 * cfg.nx * cfg.ny
    o = op.Oracle(cfg)
    o.set_state(st)
    o.step(50)
    got = o.get_state()
    for f in range(cfg.num_fields):
        assert abs(got[f, 0, 0] - st[f, 0, 0]) <= 1e-12 * max(1.0, abs(st[f, 0, 0]))
    assert np.all(got[:, ~shhd.active_mask(cfg)] == 0)
    assert np.array_equal(got, shhd.hermitian_project(got, cfg))


def test_zero_density_raises():
    cfg = shhd.SolverConfig(nx_active=8, ny_active=8, dealias="3/2", pressure_coefficient=1.0, eta=0.1,
                            momentum_advection=True)
    o = op.Oracle(cfg)
    o.set_state(np.zeros(cfg.state_shape(), complex))
    with pytest.raises(op.OracleError, match="zero density"):
        o.step(1)


def test_config_validation_messages():
    bad = shhd.SolverConfig(nx_active=8, ny_active=8, mode=shhd.MODE_QUIESCENT, momentum_advection=True)
    with pytest.raises(op.OracleError, match="quiescent time evolution cannot use momentum"):
        op.Oracle(bad)


# ---------------------------------------------------------------------------
# noise (src/fcalculator.cc:496-700)
# ---------------------------------------------------------------------------
def test_psi_noise_covariance():
    """E|out|^2 = 2 sigma^2 = P M k^2 with P = 2 kBT chi Nc^2 / V (per mode)."""
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=20.0, ly=20.0,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[0.6],
                            noise_order_parameter=True, kBT=0.3, order_parameter_noise_chi=2.0, seed=7)
    o = op.Oracle(cfg)
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    nc = cfg.nx * cfg.ny
    P = 2 * 0.3 * 2.0 * nc * nc / (20.0 * 20.0)
    mask = shhd.active_mask(cfg) & (k2 > 0)
    acc = np.zeros(mask.shape)
    ndraw = 40
    for _ in range(ndraw):
        s = o.rhs_sto()[1]
        acc += np.abs(s) ** 2
        # the kx = 0 line draws one member of each (ky, -ky) pair and writes
        # the conjugate onto the other (src/fcalculator.cc:525-566)
        for gy in range(1, cfg.ny_active // 2):
            assert s[cfg.ny - gy, 0] == np.conj(s[gy, 0])
        assert s[0, 0] == 0
    ratio = (acc / ndraw)[mask] / (P * 0.6 * k2[mask])
    assert abs(ratio.mean() - 1) < 0.02
    assert np.all(np.isfinite(ratio))


def test_j_noise_covariance():
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=20.0, ly=20.0, eta=0.3, zeta=0.7,
                            noise_momentum=True, kBT=0.5, seed=3)
    o = op.Oracle(cfg)
    kx, ky, _ = shhd.wavenumbers(cfg)
    nc = cfg.nx * cfg.ny
    s2 = 0.5 * (2 * 0.5 * nc * nc / 400.0)
    mask = shhd.active_mask(cfg) & ((kx ** 2 + ky ** 2) > 0)
    ax = np.zeros(mask.shape)
    ay = np.zeros(mask.shape)
    nd = 40
    for _ in range(nd):
        s = o.rhs_sto()
        ax += np.abs(s[1]) ** 2
        ay += np.abs(s[2]) ** 2
    ex = 2 * s2 * ((0.7 + 0.3) * kx ** 2 + 0.3 * ky ** 2) + 0 * ky
    ey = 2 * s2 * (0.3 * kx ** 2 + (0.7 + 0.3) * ky ** 2)
    assert abs(((ax / nd)[mask] / ex[mask]).mean() - 1) < 0.02
    assert abs(((ay / nd)[mask] / ey[mask]).mean() - 1) < 0.02


def psi_energy(state_psi, cfg, a, b):
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    nc = cfg.nx * cfg.ny
    w = np.full(k2.shape, 2.0)
    w[:, 0] = 1.0   # kx = 0 column holds each mode once (full-space sum)
    vol = cfg.lx * cfg.ly
    return 0.5 * np.sum(w * (a + b * k2) * np.abs(state_psi) ** 2) * vol / nc ** 2


def test_equipartition_known_answer():
    """examples/01_best_timestep_equilibrium/README.md: E_psi = kBT/2 * dof with
    dof = (Na-1)^2 - 1 (1984 at 64^2, kBT = 1).  At 32^2: 480 kBT."""
    na = 32
    a = 1.0
    cfg = shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=32.0, ly=32.0, dt=0.01,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                            free_energy_kind=shhd.FREE_ENERGY_QUADRATIC, mu_k0=[a],
                            noise_order_parameter=True, kBT=1.0, seed=11)
    rng = np.random.default_rng(0)
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    nc = cfg.nx * cfg.ny
    var = 1.0 * nc * nc / (cfg.lx * cfg.ly * a)   # E|psi_k|^2 = kBT Nc^2 / (V a)
    st = np.zeros(cfg.state_shape(), complex)
    z = (rng.normal(size=k2.shape) + 1j * rng.normal(size=k2.shape)) * math.sqrt(var / 2)
    st[1] = z
    st = shhd.hermitian_project(st, cfg)
    st[1][1:, 0] *= math.sqrt(2.0)   # kx = 0 column: projection halves the variance
    st[1][0, 0] = 0
    o = op.Oracle(cfg)
    o.set_state(st)
    es = []
    for _ in range(60):
        o.step(5)
        es.append(psi_energy(o.get_state()[1], cfg, a, 0.0))
    dof = (na - 1) ** 2 - 1
    assert abs(np.mean(es[10:]) / (0.5 * dof) - 1) < 0.04
