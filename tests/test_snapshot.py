"""Snapshot text (shhd_write_snapshot, host-only) in the reference's format
(SnapshotMeasure, src/measure_snapshot.cc:35-138): header lines, column order,
"%.16e" numbers (17 significant digits, so a spectral snapshot read back is
the state again, bit for bit) and the file-name rule."""
import numpy as np

import shhd


def small_cfg(n=1):
    return shhd.SolverConfig(nx_active=8, ny_active=8, dealias="3/2", lx=8.0, ly=8.0, num_order_parameters=n,
                             mobility=[1.0] * n)


def test_spectral_snapshot_format_and_round_trip(tmp_path):
    cfg = small_cfg()
    rng = np.random.default_rng(1)
    st = rng.normal(size=cfg.state_shape()) + 1j * rng.normal(size=cfg.state_shape())
    st[0, 0, 1] = -0.0
    p = str(tmp_path / "s.snapshot")
    shhd.write_snapshot(p, shhd.SNAPSHOT_SPECTRAL, st, cfg, 40, 0.25)
    lines = open(p).read().splitlines()
    assert lines[0] == "# snapshot space spectral"
    assert lines[1] == "# step 40 time 2.5000000000000000e-01"
    assert lines[2] == "# nkx 7 nky 12 order_parameters 1"
    assert lines[3] == ("# kx_index ky_index rho_real rho_imag psi_com0_real psi_com0_imag "
                        "jx_real jx_imag jy_real jy_imag")
    assert len(lines) == 4 + 7 * 12
    first = lines[4].split()
    assert first[:2] == ["0", "0"] and len(first) == 2 + 2 * 4
    assert first[2] == "%.16e" % st[0, 0, 0].real
    assert lines[5].split()[:2] == ["1", "0"]
    assert lines[5].split()[2] == "-0.0000000000000000e+00"
    space, step, time_, back = shhd.read_snapshot(p)
    assert (space, step, time_) == (shhd.SNAPSHOT_SPECTRAL, 40, 0.25)
    assert np.array_equal(back, st)


def test_physical_snapshot_format(tmp_path):
    cfg = small_cfg(n=2)
    phys = np.arange(cfg.num_fields * cfg.ny * cfg.nx, dtype=float).reshape(cfg.num_fields, cfg.ny, cfg.nx) / 7.0
    p = str(tmp_path / "p.snapshot")
    shhd.write_snapshot(p, shhd.SNAPSHOT_PHYSICAL, phys, cfg, 3, 1.0)
    lines = open(p).read().splitlines()
    assert lines[0] == "# snapshot space physical"
    assert lines[2] == "# nx 12 ny 12 order_parameters 2"
    assert lines[3] == "# x y rho psi_com0 psi_com1 jx jy"
    row = lines[4 + 12 + 5].split()           # gy = 1, gx = 5
    assert row[:2] == ["5", "1"]
    assert [float(v) for v in row[2:]] == [phys[f, 1, 5] for f in range(5)]
    _, _, _, back = shhd.read_snapshot(p)
    assert np.array_equal(back, phys)


def test_snapshot_filename_width():
    assert shhd.snapshot_filename("out", 100, 2000) == "out_step00100.snapshot"
    assert shhd.snapshot_filename("a/b", 7, 9) == "a/b_step07.snapshot"
