"""Host-side measures fed by the on-device probe ring (shhd_*_set_probes /
read_probes): the values the reference's measures read every step arrive in
one download per chunk of steps instead of a download of the whole flux or
state per step.

* `GreenKubo`: YokotaGreenKuboMeasure (src/measure_yokota_green_kubo.cc),
  the block integrals of Pi_xy_hat at the selected wave vectors.
* `flux_time_series_probes`: the k = 0 flux columns of TimeSeriesMeasure
  (src/measure_time_series.cc:189-227), Re F_hat(0) / (Mx*My).
* `lowest_modes` / `time_correlation`: the low-mode time correlation of an
  ensemble run.
"""
from __future__ import annotations

import math
from typing import List, Tuple

import numpy as np

import shhd


def _wrap_gy(cfg: shhd.SolverConfig, ny: int) -> int:
    return ny % cfg.ny


def green_kubo_wave_numbers(cfg: shhd.SolverConfig, selection: str = "diagonal") -> List[Tuple[int, int]]:
    """Wave numbers (nx, ny) in the reference's mode order
    (src/measure_yokota_green_kubo.cc:39-89)."""
    max_nx = max(0, cfg.nx_active // 2 - 1)
    max_ny = max(0, cfg.ny_active // 2 - 1)
    if selection == "diagonal":
        nmodes = min(max_nx, max_ny)
        out = [(n + 1, n + 1) for n in range(nmodes)]
    elif selection == "all":
        out = [(0, m + 1) for m in range(max_ny)]
        ny_count = 2 * max_ny + 1
        for s in range(max_nx * ny_count):
            out.append((s // ny_count + 1, s % ny_count - max_ny))
    else:
        raise ValueError("YokotaGreenKuboMeasure: mode must be diagonal|all.")
    if not out:
        raise ValueError("YokotaGreenKuboMeasure: no active wave vector is selected.")
    return out


def green_kubo_probes(cfg: shhd.SolverConfig, selection: str = "diagonal"):
    """Probes of Pi_xy_hat at the Green-Kubo wave vectors (flux slot 2N+3)."""
    slot = 2 + 2 * cfg.num_order_parameters + 1
    return [(shhd.PROBE_FLUX, slot, nx, _wrap_gy(cfg, ny)) for nx, ny in green_kubo_wave_numbers(cfg, selection)]


class GreenKubo:
    """YokotaGreenKuboMeasure::observe (src/measure_yokota_green_kubo.cc:128-193)
    driven by probe rows of Pi_xy_hat, one row per step."""

    def __init__(self, cfg: shhd.SolverConfig, nevery: int, nblock: int, selection: str = "diagonal"):
        if cfg.kBT <= 0.0:
            raise ValueError("YokotaGreenKuboMeasure requires positive kBT from fix noise.")
        self.cfg = cfg
        self.nevery = int(nevery)
        self.nblock = int(nblock)
        self.waves = green_kubo_wave_numbers(cfg, selection)
        self.ntau = self.nblock // self.nevery
        self.block_step = 0
        self.nsample = 0
        self.integrals = np.zeros(len(self.waves), np.complex128)
        self.sums = np.zeros((len(self.waves), self.ntau))

    def observe(self, rows: np.ndarray):
        cfg = self.cfg
        area = cfg.lx * cfg.ly
        cell_area = area / (cfg.nx * cfg.ny)
        for pi_xy in np.atleast_2d(rows):
            self.block_step += 1
            weighted_dt = 0.5 * cfg.dt if self.block_step == 1 else cfg.dt
            self.integrals += (weighted_dt * cell_area) * pi_xy
            if self.block_step % self.nevery == 0:
                tau_index = self.block_step // self.nevery - 1
                tau = self.block_step * cfg.dt
                self.sums[:, tau_index] += np.abs(self.integrals) ** 2 / (2.0 * tau * cfg.kBT * area)
            if self.block_step == self.nblock:
                self.nsample += 1
                self.block_step = 0
                self.integrals[:] = 0

    def table(self):
        """Rows (nsample, tau, mode_index, kx, ky, S_mean) as the reference's
        output table (src/measure_yokota_green_kubo.cc:195-225)."""
        cfg = self.cfg
        out = []
        for t in range(self.ntau):
            tau = (t + 1) * self.nevery * cfg.dt
            for m, (nx, ny) in enumerate(self.waves):
                kx = 2.0 * math.pi * nx / cfg.lx
                ky = 2.0 * math.pi * ny / cfg.ly
                out.append((self.nsample, tau, m, kx, ky, self.sums[m, t] / self.nsample))
        return out


def flux_time_series_probes(cfg: shhd.SolverConfig, columns):
    """Probes for TimeSeriesMeasure's flux columns: names 'pi_xx', 'pi_xy',
    'pi_yy', 'op_flux_x[a]', 'op_flux_y[a]' -> the k = 0 entry of that slot.
    The column value is probe.real / (Mx*My)."""
    n = cfg.num_order_parameters
    out = []
    for c in columns:
        if c in ("pi_xx", "pi_xy", "pi_yy"):
            slot = 2 + 2 * n + ("pi_xx", "pi_xy", "pi_yy").index(c)
        elif c.startswith("op_flux_"):
            a = int(c[c.index("[") + 1:c.index("]")])
            slot = 2 + 2 * a + (0 if c[8] == "x" else 1)
        else:
            raise ValueError(f"not a flux column: {c}")
        out.append((shhd.PROBE_FLUX, slot, 0, 0))
    return out


def lowest_modes(cfg: shhd.SolverConfig, count: int = 16):
    """The `count` active modes with the smallest |k|^2 > 0, one of each
    (k, -k) pair (kx > 0, or kx = 0 and ky > 0), as reference (gx, gy)."""
    kx, ky, sy = shhd.wavenumbers(cfg)
    mask = shhd.active_mask(cfg)
    k2 = kx ** 2 + ky ** 2
    cand = []
    for gy in range(cfg.ny):
        for gx in range(cfg.nx // 2 + 1):
            if not mask[gy, gx] or k2[gy, gx] == 0.0:
                continue
            if gx == 0 and sy[gy] < 0:
                continue
            cand.append((k2[gy, gx], gx, gy))
    cand.sort()
    return [(gx, gy) for _, gx, gy in cand[:count]]


def time_correlation(rows: np.ndarray, maxlag: int) -> np.ndarray:
    """C[lag, p] = mean over t of a_p(t + lag) * conj(a_p(t)) for a probe
    series a (rows x probes)."""
    rows = np.asarray(rows)
    nt = rows.shape[0]
    out = np.zeros((maxlag + 1, rows.shape[1]), np.complex128)
    for lag in range(maxlag + 1):
        if lag >= nt:
            break
        out[lag] = np.mean(rows[lag:] * np.conj(rows[:nt - lag]), axis=0)
    return out
