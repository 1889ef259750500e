"""Host-side measures over probe rows (measures.py), on the CPU."""
import math

import numpy as np
import pytest

import shhd
import measures


def cfg16(**kw):
    base = dict(nx_active=16, ny_active=16, dealias="3/2", lx=16.0, ly=16.0, dt=0.01, eta=0.5, zeta=0.5,
                pressure_coefficient=1.0, noise_momentum=True, kBT=0.3)
    base.update(kw)
    return shhd.SolverConfig(**base)


def test_green_kubo_constant_flux_closed_form():
    """With Pi_xy = c every step the block integral after b steps is
    c * cell * dt * (b - 1/2) (the first step carries dt/2), and each sample is
    |I|^2 / (2 tau kBT A) with tau = b dt."""
    cfg = cfg16()
    gk = measures.GreenKubo(cfg, nevery=2, nblock=6)
    c = np.full(len(gk.waves), 2.0 + 1.0j)
    gk.observe(np.tile(c, (12, 1)))          # two blocks
    area = cfg.lx * cfg.ly
    cell = area / (cfg.nx * cfg.ny)
    tab = gk.table()
    assert gk.nsample == 2
    assert len(tab) == gk.ntau * len(gk.waves)
    for nsample, tau, m, kx, ky, s in tab:
        b = round(tau / cfg.dt)
        expect = abs(c[0] * cell * cfg.dt * (b - 0.5)) ** 2 / (2.0 * tau * cfg.kBT * area)
        assert nsample == 2
        assert s == pytest.approx(expect, rel=1e-13)
        nx, ny = gk.waves[m]
        assert kx == pytest.approx(2 * math.pi * nx / cfg.lx) and ky == pytest.approx(2 * math.pi * ny / cfg.ly)


@pytest.mark.parametrize("selection", ["diagonal", "all"])
def test_green_kubo_wave_numbers_are_active_and_distinct(selection):
    cfg = cfg16(nx_active=16, ny_active=8, ly=8.0)
    waves = measures.green_kubo_wave_numbers(cfg, selection)
    max_nx, max_ny = 16 // 2 - 1, 8 // 2 - 1
    count = min(max_nx, max_ny) if selection == "diagonal" else max_ny + max_nx * (2 * max_ny + 1)
    assert len(waves) == len(set(waves)) == count
    mask = shhd.active_mask(cfg)
    for src, slot, gx, gy in measures.green_kubo_probes(cfg, selection):
        assert src == shhd.PROBE_FLUX and slot == 3 and mask[gy, gx]
    with pytest.raises(ValueError):
        measures.GreenKubo(cfg16(kBT=0.0), 1, 2)


def test_lowest_modes_order():
    cfg = cfg16()
    modes = measures.lowest_modes(cfg, 6)
    assert set(modes[:2]) == {(1, 0), (0, 1)}
    assert set(modes[2:4]) == {(1, 1), (1, cfg.ny - 1)}
    assert set(modes[4:6]) == {(2, 0), (0, 2)}


def test_time_correlation_of_a_phasor():
    w = np.array([0.3, -1.1])
    t = np.arange(50)[:, None]
    rows = 2.0 * np.exp(1j * w * t)
    c = measures.time_correlation(rows, 5)
    for lag in range(6):
        assert np.allclose(c[lag], 4.0 * np.exp(1j * w * lag), rtol=1e-14, atol=0)


def test_flux_time_series_probe_slots():
    cfg = cfg16(num_order_parameters=2, mobility=[1, 1])
    p = measures.flux_time_series_probes(cfg, ["pi_xx", "pi_xy", "pi_yy", "op_flux_x[1]", "op_flux_y[0]"])
    assert [q[1] for q in p] == [6, 7, 8, 4, 3]
    assert all(q[0] == shhd.PROBE_FLUX and q[2:] == (0, 0) for q in p)
