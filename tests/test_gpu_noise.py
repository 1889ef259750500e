"""Stochastic parity of the HIP path: the reference's mt19937_64 stream is not
reproduced (the device draws from a counter-based Philox generator), so noise
is checked in distribution against the same closed forms the oracle is pinned
to (src/fcalculator.cc:496-700, docs/equations.md:72-99) and against the
equipartition known answer of examples/01_best_timestep_equilibrium."""
import math

import numpy as np
import pytest

import shhd
import oracle_py as op
import noise_ref as nr

pytestmark = pytest.mark.gpu

S2 = math.sqrt(2.0)
S3 = math.sqrt(3.0)
B1 = (2 * S2 + S3) / 5
B2 = (-4 * S2 + 3 * S3) / 5
B3 = (S2 - 2 * S3) / 10


def banner(name):
    print(f"[case] {name}", flush=True)


def srk3_noise_gain():
    """Noise-only SRK3 step from zero: u = sqrt(dt) (cA A + cB B)."""
    cA = 1.0
    cB = (2.0 / 3.0) * ((B1 + B2) / 4.0 + B3)
    return cA * cA + cB * cB


def draw_variance(cfg, field, nrep):
    """Mean |u_k|^2 over nrep noise-only steps from the zero state."""
    acc = np.zeros(cfg.state_shape()[1:])
    zero = np.zeros(cfg.state_shape(), complex)
    zero[0, 0, 0] = cfg.nx * cfg.ny   # rho = 1 (compressible velocity needs rho != 0)
    with shhd.GpuFluidSolver(cfg) as g:
        for _ in range(nrep):
            g.set_state(zero)
            g.step(1)
            s = g.get_state()
            acc += np.abs(s[field]) ** 2
            for gy in range(1, cfg.ny_active // 2):   # kx = 0 pairs conjugate
                assert s[field, cfg.ny - gy, 0] == np.conj(s[field, gy, 0])
    return acc / nrep


@pytest.mark.parametrize("precision", [64, 32])
@pytest.mark.parametrize("integrator", [shhd.INTEGRATOR_EULER, shhd.INTEGRATOR_SRK3])
def test_psi_noise_covariance(integrator, precision):
    banner(f"psi noise covariance integrator={integrator} fp{precision}")
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=40.0, ly=40.0, dt=0.01,
                            integrator=integrator, mode=shhd.MODE_QUIESCENT, num_order_parameters=1,
                            mobility=[0.6], noise_order_parameter=True, kBT=0.3, order_parameter_noise_chi=2.0,
                            seed=7, precision=precision)
    var = draw_variance(cfg, 1, 24)
    kx, ky, _ = shhd.wavenumbers(cfg)
    k2 = kx ** 2 + ky ** 2
    nc = cfg.nx * cfg.ny
    P = 2 * 0.3 * 2.0 * nc * nc / 1600.0
    gain = srk3_noise_gain() if integrator == shhd.INTEGRATOR_SRK3 else 1.0
    mask = shhd.active_mask(cfg) & (k2 > 0)
    ratio = var[mask] / (cfg.dt * gain * P * 0.6 * k2[mask])
    print(f"    mean ratio {ratio.mean():.4f} over {mask.sum()} modes", flush=True)
    assert abs(ratio.mean() - 1) < 0.02
    # no structure left in k: low and high halves agree
    lo = ratio[k2[mask] < np.median(k2[mask])].mean()
    hi = ratio[k2[mask] >= np.median(k2[mask])].mean()
    assert abs(lo - hi) < 0.04


@pytest.mark.parametrize("precision", [64, 32])
def test_j_noise_covariance(precision):
    banner(f"momentum noise covariance fp{precision}")
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=40.0, ly=40.0, dt=0.01,
                            integrator=shhd.INTEGRATOR_EULER, eta=0.3, zeta=0.7, noise_momentum=True, kBT=0.5,
                            seed=3, precision=precision)
    vx = draw_variance(cfg, 1, 24)
    vy = draw_variance(cfg, 2, 24)
    kx, ky, _ = shhd.wavenumbers(cfg)
    nc = cfg.nx * cfg.ny
    s2 = 0.5 * (2 * 0.5 * nc * nc / 1600.0)
    mask = shhd.active_mask(cfg) & ((kx ** 2 + ky ** 2) > 0)
    ex = cfg.dt * 2 * s2 * ((0.7 + 0.3) * kx ** 2 + 0.3 * ky ** 2)
    ey = cfg.dt * 2 * s2 * (0.3 * kx ** 2 + (0.7 + 0.3) * ky ** 2)
    rx = (vx[mask] / ex[mask]).mean()
    ry = (vy[mask] / ey[mask]).mean()
    print(f"    ratios {rx:.4f} {ry:.4f}", flush=True)
    assert abs(rx - 1) < 0.02 and abs(ry - 1) < 0.02


def test_noise_deterministic_and_seeded():
    banner("noise determinism")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=32.0, ly=32.0, dt=0.01,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                            free_energy_kind=shhd.FREE_ENERGY_QUADRATIC, mu_k0=[1.0],
                            noise_order_parameter=True, kBT=1.0, seed=5)
    st = np.zeros(cfg.state_shape(), complex)
    outs = []
    for seed in (5, 5, 6):
        cfg.seed = seed
        with shhd.GpuFluidSolver(cfg) as g:
            g.set_state(st)
            g.step(3)
            outs.append(g.get_state())
    assert np.array_equal(outs[0], outs[1])
    assert not np.array_equal(outs[0], outs[2])


def psi_energy(psi, cfg, a):
    nc = cfg.nx * cfg.ny
    w = np.full(psi.shape, 2.0)
    w[:, 0] = 1.0
    return 0.5 * np.sum(w * a * np.abs(psi) ** 2) * cfg.lx * cfg.ly / nc ** 2


@pytest.mark.parametrize("precision", [64, 32])
def test_equipartition_known_answer(precision):
    """E_psi = kBT/2 ((Na-1)^2 - 1): 1984 at 64^2 and kBT = 1
    (examples/01_best_timestep_equilibrium/README.md)."""
    banner(f"equipartition 64^2 fp{precision}")
    na, a = 64, 1.0
    cfg = shhd.SolverConfig(nx_active=na, ny_active=na, dealias="3/2", lx=64.0, ly=64.0, dt=0.01,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                            free_energy_kind=shhd.FREE_ENERGY_QUADRATIC, mu_k0=[a],
                            noise_order_parameter=True, kBT=1.0, seed=11, precision=precision)
    rng = np.random.default_rng(0)
    nc = cfg.nx * cfg.ny
    var = nc * nc / (cfg.lx * cfg.ly * a)
    st = np.zeros(cfg.state_shape(), complex)
    st[1] = (rng.normal(size=st[1].shape) + 1j * rng.normal(size=st[1].shape)) * math.sqrt(var / 2)
    st = shhd.hermitian_project(st, cfg)
    st[1][1:, 0] *= math.sqrt(2.0)
    st[1][0, 0] = 0
    es = []
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(st)
        for _ in range(60):
            g.step(5)
            es.append(psi_energy(g.get_state()[1], cfg, a))
    e = float(np.mean(es[10:]))
    print(f"    E_psi = {e:.1f}, expected 1984", flush=True)
    assert abs(e / 1984.0 - 1) < 0.03


def test_oracle_and_gpu_same_noise_law():
    """The oracle (reference RNG order) and the GPU agree on the SRK3 noise
    variance per mode to sampling error."""
    banner("oracle vs gpu srk3 noise law")
    cfg = shhd.SolverConfig(nx_active=32, ny_active=32, dealias="3/2", lx=30.0, ly=30.0, dt=0.02,
                            mode=shhd.MODE_QUIESCENT, num_order_parameters=1, mobility=[1.0],
                            noise_order_parameter=True, kBT=0.7, seed=9)
    zero = np.zeros(cfg.state_shape(), complex)
    o = op.Oracle(cfg)
    acc = np.zeros(zero.shape[1:])
    n = 30
    for _ in range(n):
        o.set_state(zero)
        o.step(1)
        acc += np.abs(o.get_state()[1]) ** 2
    vo = acc / n
    vg = draw_variance(cfg, 1, n)
    mask = shhd.active_mask(cfg)
    mask[0, 0] = False
    r = vg[mask].sum() / vo[mask].sum()
    print(f"    gpu/oracle variance ratio {r:.4f}", flush=True)
    assert abs(r - 1) < 0.03


def noise_step(cfg, rho_one):
    """The state after one Euler step from zero (rho = 1 when rho_one): with
    every deterministic term zero, the increment is the step's noise alone."""
    zero = np.zeros(cfg.state_shape(), complex)
    if rho_one:
        zero[0, 0, 0] = cfg.nx * cfg.ny
    with shhd.GpuFluidSolver(cfg) as g:
        g.set_state(zero)
        g.step(1)
        return g.get_state()


@pytest.mark.parametrize("precision", [64, 32])
def test_psi_deviates_are_philox_box_muller(precision):
    """Mode by mode, the psi noise of the first step is sqrt(dt) sigma_k times
    the first Box-Muller deviate of Philox4x32-10 at counter (mode, 0, 0, 0):
    the generator, its counter layout, the seed split and the kx = 0
    conjugate pairing, against the CPU restatement pinned by its
    known-answer vectors."""
    banner(f"psi deviates fp{precision}")
    seed = 0x1_2345_6789
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=40.0, ly=40.0, dt=0.01,
                            integrator=shhd.INTEGRATOR_EULER, mode=shhd.MODE_QUIESCENT, num_order_parameters=1,
                            mobility=[0.6], noise_order_parameter=True, kBT=0.3, order_parameter_noise_chi=2.0,
                            seed=seed, precision=precision)
    got = noise_step(cfg, False)[1]
    kx, ky, _ = shhd.wavenumbers(cfg)
    nc = cfg.nx * cfg.ny
    scale = math.sqrt(cfg.dt) * np.sqrt(0.5 * (2 * 0.3 * 2.0 * nc * nc / 1600.0) * 0.6 * (kx ** 2 + ky ** 2))
    mode, conj, live = nr.mode_table(cfg)
    z0, _ = nr.gauss_quad(mode, 0, 0, seed)
    want = scale * np.where(conj, np.conj(z0), z0)
    err = np.abs(got - want)[live] / scale[live]
    print(f"    {live.sum()} modes, max deviate error {err.max():.2e}", flush=True)
    assert err.max() < 1e-4
    assert np.all(got[~live] == 0)


@pytest.mark.parametrize("precision", [64, 32])
def test_momentum_deviates_are_philox_box_muller(precision):
    """The momentum noise of the first Euler step from rest, mode by mode:
    the stress deviates x1, x2 (both pairs of stream 8) and x3 (first pair of
    stream 10) combined into -i k . Pi with Pi_xx, Pi_yy = -s (sqrt(zeta) x1
    +- sqrt(eta) x2), Pi_xy = -s sqrt(eta) x3 (src/fcalculator.cc:586-700)."""
    banner(f"momentum deviates fp{precision}")
    seed = 77
    cfg = shhd.SolverConfig(nx_active=64, ny_active=64, dealias="3/2", lx=40.0, ly=40.0, dt=0.01,
                            integrator=shhd.INTEGRATOR_EULER, eta=0.3, zeta=0.7, noise_momentum=True, kBT=0.5,
                            seed=seed, precision=precision)
    s = noise_step(cfg, True)
    kx, ky, _ = shhd.wavenumbers(cfg)
    kx = np.broadcast_to(kx, s[1].shape)
    ky = np.broadcast_to(ky, s[1].shape)
    nc = cfg.nx * cfg.ny
    ss = math.sqrt(0.5 * 2 * 0.5 * nc * nc / 1600.0)
    se, sz = math.sqrt(0.3), math.sqrt(0.7)
    mode, conj, live = nr.mode_table(cfg)
    x1, x2 = nr.gauss_quad(mode, 8, 0, seed)
    x3, _ = nr.gauss_quad(mode, 10, 0, seed)
    kxd = np.where(conj, 0.0, kx)
    kyd = np.where(conj, -ky, ky)
    pxx = -ss * (sz * x1 + se * x2)
    pyy = -ss * (sz * x1 - se * x2)
    pxy = -ss * se * x3
    vx = -1j * (kxd * pxx + kyd * pxy)
    vy = -1j * (kxd * pxy + kyd * pyy)
    vx = np.where(conj, np.conj(vx), vx)
    vy = np.where(conj, np.conj(vy), vy)
    scale = math.sqrt(cfg.dt) * ss * np.sqrt(kx ** 2 + ky ** 2)
    for f, v in ((1, vx), (2, vy)):
        err = np.abs(s[f] - math.sqrt(cfg.dt) * v)[live] / scale[live]
        print(f"    field {f}: max deviate error {err.max():.2e}", flush=True)
        assert err.max() < 1e-4
        assert np.all(s[f][~live] == 0)
